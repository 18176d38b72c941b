"""FP64 NumPy solutions that batched and served results report their error against.

Reporting only: these never produce a result, they measure one. The error of each result is
stated against a float64 solution of the same problem (the accuracy statement the reference's
validation makes against closed forms, ``src/pyheat3d/validation.py:36-147``, extended to every
request of the mixed workloads):

* heat: the same initial field advanced by the same number of FTCS steps in float64 (the
  reference's update order, ``src/pyheat3d/backends/numpy_ref.py:30-47``, with the float32
  coefficients the device uses);
* Jacobi Laplace / Poisson: the exact solution of the discrete Dirichlet system, from a sparse
  direct solve (SciPy) in float64; Jacobi converges to it, so its distance measures both the
  residual left and the storage rounding.

Errors are the reference's max-norm relative deviation ``max|u - ref| / max|ref|``
(``validation.py:20-21``).
"""

from __future__ import annotations

import numpy as np


def rel_err(u: np.ndarray, ref: np.ndarray) -> float:
    u = np.asarray(u, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(u - ref)) / max(float(np.max(np.abs(ref))), 1e-300))


def heat(u0: np.ndarray, c0: float, c1: float, bc: str, steps: int) -> np.ndarray:
    """``steps`` FTCS steps in float64 (neighbour pairs summed x first, then y, then z)."""
    u = np.asarray(u0, dtype=np.float64).copy()
    c0, c1 = float(np.float32(c0)), float(np.float32(c1))
    nd = u.ndim
    if bc == "periodic":
        for _ in range(steps):
            s = None
            for ax in range(nd - 1, -1, -1):
                pair = np.roll(u, 1, ax) + np.roll(u, -1, ax)
                s = pair if s is None else s + pair
            u = c0 * u + c1 * s
        return u
    core = (slice(1, -1),) * nd
    out = u.copy()
    for _ in range(steps):
        s = None
        for ax in range(nd - 1, -1, -1):
            lo = [slice(1, -1)] * nd
            hi = [slice(1, -1)] * nd
            lo[ax], hi[ax] = slice(0, -2), slice(2, None)
            pair = u[tuple(lo)] + u[tuple(hi)]
            s = pair if s is None else s + pair
        out[core] = c0 * u[core] + c1 * s
        u, out = out, u
    return u


def dirichlet_solution(u0: np.ndarray, c1: float, source: np.ndarray | None = None) -> np.ndarray:
    """Fixed point of the Jacobi map ``u = c1 * S(u) + b`` inside the box in float64, with the
    outer layer of ``u0`` as Dirichlet data (``b`` = the source, 0 for Laplace).

    ``c1`` is the device's float32 coefficient widened (float32(1/6) differs from 1/6 by 3e-8,
    which moves the 3D fixed point by up to ~1e-4 relative at 256^3), so the error against this
    solution measures the rounding and the residual left, not the coefficient.

    ``(I - c1 * N) u = c1 * S(boundary) + b`` with N the interior neighbour-sum operator: N is a
    sum of tridiagonal (0, 1, 1) matrices, one per axis, each diagonalised by the type-I
    discrete sine transform with eigenvalues 2 cos(pi k / (m + 1)), k = 1..m. So the solve is a
    forward DST, a division by 1 - c1 * sum of those eigenvalues, and the inverse DST.
    """
    from scipy.fft import dstn, idstn

    g = np.asarray(u0, dtype=np.float64)
    c1 = float(np.float32(c1))
    nd = g.ndim
    inner = tuple(n - 2 for n in g.shape)
    core = (slice(1, -1),) * nd
    data = g.copy()
    data[core] = 0.0   # the neighbour sum of the boundary data alone
    rhs = np.zeros(inner)
    for ax in range(nd):
        lo = [slice(1, -1)] * nd
        hi = [slice(1, -1)] * nd
        lo[ax], hi[ax] = slice(0, -2), slice(2, None)
        rhs += data[tuple(lo)] + data[tuple(hi)]
    rhs *= c1
    if source is not None:
        rhs += np.asarray(source, dtype=np.float64)[core]
    lam = np.ones(inner)
    for ax, m in enumerate(inner):
        shape = [1] * nd
        shape[ax] = m
        k = np.arange(1, m + 1, dtype=np.float64).reshape(shape)
        lam = lam - c1 * 2.0 * np.cos(np.pi * k / (m + 1))
    out = g.copy()
    out[core] = idstn(dstn(rhs, type=1, norm="ortho") / lam, type=1, norm="ortho")
    return out
