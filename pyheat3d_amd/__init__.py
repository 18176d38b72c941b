"""pyheat3d_amd: the pyheat3d stencil hot path as hand-written HIP kernels for MI355X (gfx950).

Public API mirrors the reference package (``src/pyheat3d/__init__.py``): ``Stencil``,
``heat_stencil``, ``jacobi_stencil``, ``stability_limit``, ``VARIANTS``, ``available_variants``
and ``make_stepper``; the variants are the ``hip-cdna4/*`` family.
"""

from .problem import Stencil, heat_stencil, jacobi_stencil, stability_limit
from .registry import VARIANTS, available_variants, make_stepper

__version__ = "0.1.0"

__all__ = [
    "VARIANTS",
    "Stencil",
    "available_variants",
    "heat_stencil",
    "jacobi_stencil",
    "make_stepper",
    "stability_limit",
]
