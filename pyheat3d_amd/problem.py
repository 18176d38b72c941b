"""The stencil update every variant computes, and its heat / Jacobi coefficients.

One update is applied to ``u[y, x]`` (2D) or ``u[z, y, x]`` (3D), C order, x contiguous::

    u'[c] = c0 * u[c] + c1 * S[c]  (+ b[c])
    S     = ((u[x-1] + u[x+1]) + (u[y-1] + u[y+1])) + (u[z-1] + u[z+1])

The bracketing is part of the contract: every kernel adds the x pair, then the y pair, then the
z pair, rounds ``c0*u`` and ``c1*S`` separately, adds them, and adds ``b`` last. With FMA
contraction disabled this reproduces the NumPy oracle bit for bit in float32.
(Reference: ``src/pyheat3d/problem.py:1-22``.)

Heat mode (explicit FTCS): ``c0 = 1 - 2 d r``, ``c1 = r`` with ``r = alpha dt / h^2``; stable for
``r <= 1/(2d)``.  Jacobi mode: ``c0 = 0``, ``c1 = 1/(2d)``, optional source ``b``.

Boundary conditions: ``"dirichlet"`` keeps the outer layer fixed (never written);
``"periodic"`` updates every site with wrapped neighbours.
"""

from __future__ import annotations

from dataclasses import dataclass

BCS = ("dirichlet", "periodic")
MODES = ("heat", "jacobi")

#: slack added to the von Neumann limit so that r = 1/(2d) computed in floating point passes
_LIMIT_SLACK = 1e-15


@dataclass(frozen=True)
class Stencil:
    """``u' = c0 u + c1 S`` with a boundary condition (reference ``problem.py:33-43``)."""

    c0: float
    c1: float
    bc: str = "dirichlet"

    def __post_init__(self) -> None:
        if self.bc not in BCS:
            raise ValueError(f"bc must be one of {BCS}, got {self.bc!r}")


def stability_limit(dims: int) -> float:
    """Largest stable FTCS mesh Fourier number in ``dims`` dimensions: ``1 / (2 dims)``."""
    return 0.5 / dims


def _check_dims(dims: int) -> None:
    if dims not in (2, 3):
        raise ValueError("dims must be 2 or 3")


def heat_stencil(dims: int, r: float, bc: str = "dirichlet", check: bool = True) -> Stencil:
    """FTCS heat stencil for mesh Fourier number ``r`` (reference ``problem.py:51-62``).

    Raises ``ValueError`` when ``r`` is outside ``(0, 1/(2 dims)]``.
    """
    _check_dims(dims)
    limit = stability_limit(dims)
    if check and not (0.0 < r <= limit + _LIMIT_SLACK):
        raise ValueError(f"r = {r:g} violates the FTCS stability limit r <= 1/(2d) = {limit:g}")
    return Stencil(c0=1.0 - 2.0 * dims * r, c1=r, bc=bc)


def jacobi_stencil(dims: int, bc: str = "dirichlet") -> Stencil:
    """Jacobi sweep for Laplace/Poisson: ``c0 = 0``, ``c1 = 1/(2 dims)`` (``problem.py:65-69``)."""
    _check_dims(dims)
    return Stencil(c0=0.0, c1=0.5 / dims, bc=bc)


def dt_from_r(r: float, h: float, alpha: float = 1.0) -> float:
    """Time step of mesh Fourier number ``r`` on spacing ``h``."""
    return r * h * h / alpha
