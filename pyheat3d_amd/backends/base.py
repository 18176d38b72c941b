"""The interface every variant implements (reference ``src/pyheat3d/backends/base.py:12-74``)."""

from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np

from ..problem import Stencil


class Stepper(ABC):
    """Owns a ping-pong pair of device buffers and advances the stencil update.

    ``u0`` has shape ``(ny, nx)`` or ``(nz, ny, nx)``; for Dirichlet problems its outer layer
    holds the fixed boundary values. ``source`` is an optional additive term ``b`` of the same
    shape (only for variants registered with ``source=True``).
    """

    #: device storage type ("float32", "float16", "bfloat16")
    storage: str = "float32"
    #: time steps fused per memory pass
    steps_per_pass: int = 1

    def __init__(self, u0: np.ndarray, stencil: Stencil, source: np.ndarray | None = None):
        if u0.ndim not in (2, 3):
            raise ValueError("u0 must be 2D or 3D")
        if min(u0.shape) < 3:
            raise ValueError("every axis needs at least 3 points")
        if source is not None and tuple(source.shape) != tuple(u0.shape):
            raise ValueError(f"source shape {source.shape} != field shape {u0.shape}")
        self.shape: tuple[int, ...] = tuple(int(s) for s in u0.shape)
        self.dims = u0.ndim
        self.stencil = stencil
        self.c0 = float(stencil.c0)
        self.c1 = float(stencil.c1)
        self.bc = stencil.bc
        self.periodic = stencil.bc == "periodic"
        self.has_source = source is not None
        self.steps_done = 0

    @property
    def ncells(self) -> int:
        """Sites updated per step (interior only for Dirichlet)."""
        if self.periodic:
            return int(np.prod(self.shape))
        return int(np.prod([s - 2 for s in self.shape]))

    def step(self, n: int = 1) -> None:
        """Enqueue ``n`` time steps / Jacobi sweeps (asynchronous on a device)."""
        for _ in range(n):
            self._step_once()
        self.steps_done += n

    def _step_once(self) -> None:  # pragma: no cover - overridden
        raise NotImplementedError

    @abstractmethod
    def get(self) -> np.ndarray:
        """Current state as a fresh host float32 array."""

    def synchronize(self) -> None:  # noqa: B027
        """Block until all enqueued work is done."""

    def close(self) -> None:  # noqa: B027
        """Release device memory."""
