"""Closed-form fields and decay factors on the unit square / cube, and the workload initial states.

Axis convention (reference ``src/pyheat3d/analytic.py:1-16``): a Dirichlet axis has ``n`` points
including both boundary points, ``h = 1/(n-1)``; a periodic axis has ``n`` points per period,
``h = 1/n``. Arrays are ``(y, x)`` or ``(z, y, x)``.
"""

from __future__ import annotations

import math

import numpy as np


def grid_spacing(n: int, bc: str) -> float:
    """Mesh width of an ``n``-point axis on [0, 1] (``analytic.py:14-16``)."""
    return 1.0 / (n - 1) if bc == "dirichlet" else 1.0 / n


def axis(n: int, bc: str) -> np.ndarray:
    """Float64 coordinates ``i * h`` of an ``n``-point axis."""
    return np.arange(n, dtype=np.float64) * grid_spacing(n, bc)


def mesh(n: int, dims: int, bc: str) -> tuple[np.ndarray, ...]:
    """Coordinate arrays in array order (``ij`` indexing)."""
    a = axis(n, bc)
    return tuple(np.meshgrid(*([a] * dims), indexing="ij"))


def wavenumber(m: int, bc: str) -> float:
    """``m pi`` for Dirichlet, ``2 pi m`` for periodic axes (``analytic.py:30-32``)."""
    return (m if bc == "dirichlet" else 2.0 * m) * math.pi


def _sine_axis(n: int, bc: str, m: int) -> np.ndarray:
    s = np.sin(wavenumber(m, bc) * axis(n, bc))
    if bc == "dirichlet":
        s[[0, -1]] = 0.0  # sin(m pi) is not exactly zero in floating point
    return s


def fourier_mode(n: int, dims: int, bc: str, m: int = 1) -> np.ndarray:
    """``prod_i sin(k x_i)`` (float64), exactly zero on Dirichlet walls (``analytic.py:35-45``)."""
    s = _sine_axis(n, bc, m)
    out = s
    for _ in range(dims - 1):
        out = np.multiply.outer(out, s)
    return out


def decay_continuous(t: float, dims: int, bc: str, m: int = 1, alpha: float = 1.0) -> float:
    """Mode amplitude of ``u_t = alpha lap u`` at time ``t``: ``exp(-alpha d k^2 t)``."""
    k = wavenumber(m, bc)
    return math.exp(-alpha * dims * k * k * t)


def decay_semidiscrete(t: float, n: int, dims: int, bc: str, m: int = 1,
                       alpha: float = 1.0) -> float:
    """Amplitude under the exact-in-time central-difference semi-discretisation."""
    k = wavenumber(m, bc)
    h = grid_spacing(n, bc)
    lam = 4.0 * dims * math.sin(0.5 * k * h) ** 2 / (h * h)
    return math.exp(-alpha * lam * t)


def ftcs_amplification(r: float, n: int, dims: int, bc: str, m: int = 1) -> float:
    """Per-step FTCS factor ``g = 1 - 4 d r sin^2(k h / 2)`` of the mode (``analytic.py:64-68``)."""
    k = wavenumber(m, bc)
    h = grid_spacing(n, bc)
    return 1.0 - 4.0 * dims * r * math.sin(0.5 * k * h) ** 2


def harmonic_sinh(n: int, dims: int) -> np.ndarray:
    """Harmonic field for the Laplace test (``analytic.py:71-85``).

    2D ``sin(pi x) sinh(pi y) / sinh(pi)``; 3D ``sin(pi x) sin(pi y) sinh(sqrt2 pi z)/sinh(sqrt2 pi)``.
    """
    pts = mesh(n, dims, "dirichlet")
    x = pts[-1]
    if dims == 2:
        y = pts[0]
        return np.sin(math.pi * x) * np.sinh(math.pi * y) / math.sinh(math.pi)
    z, y = pts[0], pts[1]
    q = math.pi * math.sqrt(2.0)
    return np.sin(math.pi * x) * np.sin(math.pi * y) * np.sinh(q * z) / math.sinh(q)


def harmonic_quadratic(n: int, dims: int) -> np.ndarray:
    """``x^2 - y^2`` (2D) or ``x^2 + y^2 - 2 z^2`` (3D): the 5/7-point Laplacian is exact on it."""
    pts = mesh(n, dims, "dirichlet")
    x, y = pts[-1], pts[-2]
    if dims == 2:
        return x * x - y * y
    z = pts[0]
    return x * x + y * y - 2.0 * z * z


def poisson_manufactured(n: int, dims: int) -> tuple[np.ndarray, np.ndarray]:
    """``(u, f)`` with ``u = prod sin(pi x_i)`` and ``f = lap u = -d pi^2 u`` (``analytic.py:101-108``)."""
    u = fourier_mode(n, dims, "dirichlet", 1)
    return u, -dims * math.pi ** 2 * u


def interior(a: np.ndarray) -> np.ndarray:
    """View without the outermost layer on every axis."""
    return a[(slice(1, -1),) * a.ndim]


# ----------------------------------------------------------------------------------------------
# Workload initial states
# ----------------------------------------------------------------------------------------------

def random_dirichlet(shape: tuple[int, ...], seed: int = 0) -> np.ndarray:
    """Interior ``U[0,1)`` float32 from ``default_rng(seed).random(shape, float32)``, walls 0.

    The whole array is drawn (so the stream does not depend on the wall thickness) and the outer
    layer is then set to zero.
    """
    u = np.random.default_rng(seed).random(shape, dtype=np.float32)
    for ax in range(u.ndim):
        idx = [slice(None)] * u.ndim
        idx[ax] = [0, -1]
        u[tuple(idx)] = 0.0
    return u


def sine_modes_periodic(n: int, dims: int = 3, modes: tuple[int, ...] = (1, 2, 3),
                        noise: float = 1e-3, seed: int = 1) -> np.ndarray:
    """``sum_m prod_axes sin(2 pi m x_i) + noise * N(0,1)`` on the periodic grid, as float32.

    The modes are summed in float64 and rounded once; the noise is drawn in float32 from
    ``default_rng(seed).standard_normal(n**dims, float32)`` and added in float32. Built plane by
    plane so a 1024^3 field needs no float64 temporary of full size.
    """
    shape = (n,) * dims
    out = np.empty(shape, dtype=np.float32)
    axes = [_sine_axis(n, "periodic", m) for m in modes]
    rng = np.random.default_rng(seed)
    plane = n ** (dims - 1)
    flat = out.reshape(n, plane)
    xy = [np.multiply.outer(s, s) if dims == 3 else s for s in axes]
    for i in range(n):
        row = np.zeros(xy[0].shape, dtype=np.float64)
        for s, sxy in zip(axes, xy):
            row += s[i] * sxy
        flat[i] = row.reshape(-1).astype(np.float32)
    if noise:
        for i in range(n):
            flat[i] += np.float32(noise) * rng.standard_normal(plane, dtype=np.float32)
    return out


def poisson_eigen_source(n: int, dims: int = 3) -> tuple[np.ndarray, np.ndarray, float]:
    """Manufactured Poisson problem ``-lap u = f``, ``u* = prod sin(pi x_i)``, ``f = d pi^2 u*``.

    Returns ``(v, b, alpha)`` in float64: ``v`` is the discrete eigenmode (= ``u*`` sampled),
    ``b = h^2 f / (2 d)`` is the Jacobi source, and ``alpha = d pi^2 / lambda_h`` with
    ``lambda_h = 4 d sin^2(pi h/2) / h^2`` is the amplitude of the exact discrete solution
    ``alpha * v`` (the FP64 solution of the 7-point system).
    """
    h = 1.0 / (n - 1)
    v = fourier_mode(n, dims, "dirichlet", 1)
    b = (h * h * dims * math.pi ** 2 / (2 * dims)) * v
    lam = 4.0 * dims * math.sin(0.5 * math.pi * h) ** 2 / (h * h)
    return v, b, dims * math.pi ** 2 / lam


def jacobi_eigen_iterate(n: int, sweeps: int, dims: int = 3) -> float:
    """Amplitude of the Jacobi iterate on the eigenmode problem after ``sweeps`` from zero.

    ``u_s = alpha (1 - cos(pi h)^s) v`` because the Jacobi operator maps ``v`` to ``cos(pi h) v``
    and the source to itself; the relative residual is exactly ``cos(pi h)^s``.
    """
    h = 1.0 / (n - 1)
    lam = 4.0 * dims * math.sin(0.5 * math.pi * h) ** 2 / (h * h)
    alpha = dims * math.pi ** 2 / lam
    return alpha * (1.0 - math.cos(math.pi * h) ** sweeps)


_M64 = (1 << 64) - 1


def splitmix_field(shape: tuple[int, ...], seed: int, kind: int = 2) -> np.ndarray:
    """Host mirror of ``ph3_fill_random``: the same float32 field, bit for bit.

    Site ``i`` (C order) gets ``(splitmix64(seed * 0xD1B54A32D192ED03 + i) >> 40) * 2^-24``.
    ``kind`` 0: random inside and 0 on the outer layer (heat initial state); 1: random outer
    layer and 0 inside (Laplace boundary data, zero initial guess); 2: random everywhere.
    """
    n = int(np.prod(shape))
    z = np.arange(n, dtype=np.uint64) + np.uint64((seed * 0xD1B54A32D192ED03) & _M64)
    z += np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    v = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
    v = v.reshape(shape)
    if kind in (0, 1):
        edge = np.zeros(shape, dtype=bool)
        for ax in range(len(shape)):
            idx = [slice(None)] * len(shape)
            idx[ax] = [0, -1]
            edge[tuple(idx)] = True
        v[edge if kind == 0 else ~edge] = 0.0
    return v
