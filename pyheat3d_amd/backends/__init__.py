"""Stepper implementations. ``hip_backend`` is the only compute backend of this package."""
