"""NumPy oracle: the reference's ``slicing`` backend and storage-rounding reference, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by ``tests/``, ``smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; never by the ``pyheat3d_amd`` product path.

How each function is pinned (checks live in ``tests/test_oracle.py``):

* :func:`reference_step`, :class:`NumpyStepper` -- bitwise against the scalar C restatement
  ``oracle/stencil_ref.c`` (one correctly rounded IEEE op at a time in the canonical order) on
  ragged shapes, both BCs, fp32 and fp64, with and without a source; the FTCS sine mode is
  reproduced to fp64 round-off (closed form ``g**n``); the Dirichlet boundary is untouched;
  periodic heat conserves the mean; maximum principle; a planted swap of two summation pairs or
  a dropped term is detected by the C comparison.
* :func:`round_storage` -- bit-pattern known answers for binary16 / bfloat16 round-to-nearest-
  even (ties, overflow, subnormals) and an exhaustive comparison with the C codec on a dense
  sweep of float32 bit patterns.
* :func:`emulated_reference` -- rounding cadence checked on a case where fp32 and storage
  results must differ, and k=1 equivalence with the per-step rounding.
* :func:`window_reference` -- equals the same window of the full-grid oracle bit for bit in 2D
  and 3D, both BCs, windows at corners, faces and across the periodic wrap; one layer of margin
  too few is detected.
* validation studies (:func:`fourier_run`, :func:`spatial_convergence`,
  :func:`temporal_convergence`, :func:`laplace_convergence`, :func:`laplace_quadratic`,
  :func:`poisson_convergence`) -- against the FP64 numbers the reference committed in
  ``results/validation.json`` (iteration counts exactly, errors to 1e-9 relative).
* :func:`residual_fp64` -- closed form on the Jacobi eigenmode problem (relative residual
  ``cos(pi h)**s``) and against the direct definition in float64 on random inputs.
"""

from __future__ import annotations

import math
from typing import Any

import numpy as np


# ----------------------------------------------------------------------------------------------
# the update (reference src/pyheat3d/backends/numpy_ref.py:11-47)
# ----------------------------------------------------------------------------------------------

def shift_pairs(u: np.ndarray, periodic: bool) -> list[tuple[np.ndarray, np.ndarray]]:
    """(minus, plus) neighbour arrays for x, then y, then z (``numpy_ref.py:11-27``).

    Dirichlet: views over the interior. Periodic: wrapped copies of the whole array.
    """
    d = u.ndim
    out = []
    for ax in reversed(range(d)):
        if periodic:
            out.append((np.roll(u, 1, axis=ax), np.roll(u, -1, axis=ax)))
            continue
        lo = [slice(1, -1)] * d
        hi = [slice(1, -1)] * d
        lo[ax], hi[ax] = slice(None, -2), slice(2, None)
        out.append((u[tuple(lo)], u[tuple(hi)]))
    return out


def reference_step(u: np.ndarray, c0: float, c1: float, bc: str,
                   b: np.ndarray | None = None) -> np.ndarray:
    """One update in the canonical order, as a new array (``numpy_ref.py:30-47``).

    ``s = (xm+xp) + (ym+yp) [+ (zm+zp)]``; ``new = T(c0)*u + T(c1)*s``; ``new += b``.
    """
    periodic = bc == "periodic"
    t = u.dtype.type
    pairs = shift_pairs(u, periodic)
    s = pairs[0][0] + pairs[0][1]
    for lo, hi in pairs[1:]:
        s = s + (lo + hi)
    core = (slice(None) if periodic else slice(1, -1),) * u.ndim
    new = t(c0) * u[core] + t(c1) * s
    if b is not None:
        new = new + b[core]
    out = u.copy()
    out[core] = new
    return out


class NumpyStepper:
    """Allocation-free slicing stepper (``numpy_ref.py:50-92``); the timed CPU baseline.

    Same rounding sequence as :func:`reference_step`:
    ``t1 = xm+xp; t2 = ym+yp; t1 += t2; [t2 = zm+zp; t1 += t2]; t1 *= c1; o = u*c0; o += t1;
    [o += b]``. Periodic steps delegate to :func:`reference_step`.
    """

    def __init__(self, u0: np.ndarray, c0: float, c1: float, bc: str,
                 source: np.ndarray | None = None):
        dtype = u0.dtype if u0.dtype in (np.float32, np.float64) else np.dtype(np.float32)
        self.a = np.array(u0, dtype=dtype, copy=True)
        self.b = self.a.copy()
        self.src = None if source is None else np.asarray(source, dtype=dtype)
        self.c0, self.c1, self.bc = c0, c1, bc
        self._c0 = np.dtype(dtype).type(c0)
        self._c1 = np.dtype(dtype).type(c1)
        self.periodic = bc == "periodic"
        self.steps_done = 0
        if not self.periodic:
            core = tuple(s - 2 for s in u0.shape)
            self.t1 = np.empty(core, dtype)
            self.t2 = np.empty(core, dtype)

    def _once(self) -> None:
        u, o = self.a, self.b
        if self.periodic:
            o[...] = reference_step(u, self.c0, self.c1, "periodic", self.src)
        else:
            c = (slice(1, -1),) * u.ndim
            pairs = shift_pairs(u, False)
            np.add(*pairs[0], out=self.t1)
            for lo, hi in pairs[1:]:
                np.add(lo, hi, out=self.t2)
                self.t1 += self.t2
            self.t1 *= self._c1
            oc = o[c]
            np.multiply(u[c], self._c0, out=oc)
            oc += self.t1
            if self.src is not None:
                oc += self.src[c]
        self.a, self.b = o, u

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self._once()
        self.steps_done += n

    def get(self) -> np.ndarray:
        return self.a.copy()


# ----------------------------------------------------------------------------------------------
# storage rounding (reference src/pyheat3d/validation.py:150-183, backends/cupy_backend.py:103-112)
# ----------------------------------------------------------------------------------------------

#: max relative deviation per storage type (reference ``validation.py:153-154``)
TOLERANCE = {"float32": 2e-6, "float16": 4e-3, "bfloat16": 3e-2, "native": 0.0,
             "float64": 1e-12}


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit pattern, round to nearest even on the uint32 bits."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits: np.ndarray) -> np.ndarray:
    """bfloat16 bit pattern -> float32 (exact)."""
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def round_storage(u: np.ndarray, storage: str) -> np.ndarray:
    """Round float32 values to the storage type and back (``validation.py:157-165``)."""
    if storage == "float16":
        return np.asarray(u, dtype=np.float32).astype(np.float16).astype(np.float32)
    if storage == "bfloat16":
        return bf16_value(bf16_bits(u))
    return u


def emulated_reference(u0: np.ndarray, c0: float, c1: float, bc: str, steps: int,
                       storage: str = "float32", k: int = 1,
                       b: np.ndarray | None = None) -> np.ndarray:
    """fp32 oracle with the storage rounding of a ``k``-step-per-pass variant.

    Values are rounded to storage after each pass of ``k`` fused steps; the ``steps % k``
    remainder steps are rounded one by one (``validation.py:168-183``). ``b`` (float32) is the
    optional source, added after ``c0*u + c1*S`` in every step.
    """
    u = round_storage(np.asarray(u0, dtype=np.float32), storage)
    bb = None if b is None else np.asarray(b, dtype=np.float32)
    passes, rem = divmod(steps, k)
    for _ in range(passes):
        for _ in range(k):
            u = reference_step(u, c0, c1, bc, bb)
        u = round_storage(u, storage)
    for _ in range(rem):
        u = round_storage(reference_step(u, c0, c1, bc, bb), storage)
    return u


def window_reference(u0: np.ndarray, lo: tuple[int, ...], size: int, c0: float, c1: float,
                     bc: str, steps: int, storage: str = "float32", k: int = 1,
                     b: np.ndarray | None = None, margin: int | None = None,
                     engine: str = "numpy") -> tuple[np.ndarray, list[np.ndarray]]:
    """:func:`emulated_reference` on one cubic window of a field, without the full grid.

    A site after t steps depends only on the sites within distance t of it (the light cone of
    the 2d+1-point stencil; storage rounding and the source act site by site), so the window
    padded by ``margin`` = t sites on each side (wrapped around for periodic fields, clipped at
    the fixed Dirichlet walls) and stepped with its outer layer held fixed (``"dirichlet"``) is
    exact from t layers in. ``lo`` is the window's first index per axis (periodic windows may
    run past the end and wrap). Returns the window and, per axis, the indices of its sites in
    the full field. ``engine="c"`` steps float32 windows with the scalar C loop
    (``oracle/cref.py``, bitwise equal to :func:`reference_step` and about 4x faster).
    """
    n = u0.shape
    m = steps if margin is None else margin

    def emulated(u, b_):
        if engine == "c" and storage == "float32":
            from . import cref
            return cref.run_f32(u, c0, c1, "dirichlet", steps, b_)
        return emulated_reference(u, c0, c1, "dirichlet", steps, storage, k, b_)

    if bc == "periodic":
        pad = np.ix_(*[np.arange(a - m, a + size + m) % na for a, na in zip(lo, n)])
        ref = emulated(u0[pad], None if b is None else b[pad])
        core = tuple(slice(m, m + size) for _ in n)
        return ref[core], [np.arange(a, a + size) % na for a, na in zip(lo, n)]
    if any(a < 0 or a + size > na for a, na in zip(lo, n)):
        raise ValueError("a Dirichlet window lies inside the grid")
    start = [max(a - m, 0) for a in lo]
    stop = [min(a + size + m, na) for a, na in zip(lo, n)]
    box = tuple(slice(a, e) for a, e in zip(start, stop))
    ref = emulated(u0[box], None if b is None else b[box])
    core = tuple(slice(a - s, a - s + size) for a, s in zip(lo, start))
    return ref[core], [np.arange(a, a + size) for a in lo]


def rel_err(u: np.ndarray, ref: np.ndarray) -> float:
    """Max-norm relative deviation ``max|u-ref| / max|ref|`` (``validation.py:20-21``)."""
    u = np.asarray(u, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(u - ref)) / max(np.max(np.abs(ref)), 1e-300))


# ----------------------------------------------------------------------------------------------
# Jacobi residual (new function; no counterpart in the reference, which runs fixed sweep counts)
# ----------------------------------------------------------------------------------------------

def residual_fp64(u: np.ndarray, c1: float, bc: str, b: np.ndarray | None = None,
                  c0: float = 0.0) -> float:
    """Squared 2-norm of ``b + c0*u + c1*S(u) - u`` over the updated sites, in float64.

    Every stored value is widened to float64 first; ``c0``/``c1`` are the float32 coefficients
    the kernels use, widened. This is the defect of one Jacobi sweep, proportional to
    ``f - A u`` (factor ``h^2 / (2d)``).
    """
    u64 = np.asarray(u, dtype=np.float64)
    c0f = float(np.float32(c0))
    c1f = float(np.float32(c1))
    periodic = bc == "periodic"
    pairs = shift_pairs(u64, periodic)
    s = pairs[0][0] + pairs[0][1]
    for lo, hi in pairs[1:]:
        s = s + (lo + hi)
    core = (slice(None) if periodic else slice(1, -1),) * u.ndim
    r = c0f * u64[core] + c1f * s - u64[core]
    if b is not None:
        r = r + np.asarray(b, dtype=np.float64)[core]
    return float(np.sum(r * r))


# ----------------------------------------------------------------------------------------------
# FP64 validation studies (reference src/pyheat3d/validation.py:36-147), NumPy engine only
# ----------------------------------------------------------------------------------------------

def _axis_sine(n: int, bc: str, m: int = 1) -> np.ndarray:
    k = (m if bc == "dirichlet" else 2 * m) * math.pi
    h = 1.0 / (n - 1) if bc == "dirichlet" else 1.0 / n
    s = np.sin(k * (np.arange(n, dtype=np.float64) * h))
    if bc == "dirichlet":
        s[0] = s[-1] = 0.0
    return s


def sine_mode(n: int, dims: int, bc: str, m: int = 1) -> np.ndarray:
    """``prod sin(k x_i)`` (``analytic.py:35-45``)."""
    s = _axis_sine(n, bc, m)
    out = s
    for _ in range(dims - 1):
        out = np.multiply.outer(out, s)
    return out


def fourier_run(n: int, dims: int, bc: str, r: float, t_final: float,
                dtype: type = np.float64, alpha: float = 1.0) -> dict[str, Any]:
    """Decay of the m=1 sine mode vs three references (``validation.py:36-64``)."""
    h = 1.0 / (n - 1) if bc == "dirichlet" else 1.0 / n
    steps = max(1, int(math.ceil(t_final / (r * h * h / alpha))))
    dt = t_final / steps
    r_eff = alpha * dt / (h * h)
    k = (1 if bc == "dirichlet" else 2) * math.pi
    shape0 = sine_mode(n, dims, bc)
    st = NumpyStepper(shape0.astype(dtype), 1.0 - 2.0 * dims * r_eff, r_eff, bc)
    st.step(steps)
    u = st.get().astype(np.float64)
    amp_c = math.exp(-alpha * dims * k * k * t_final)
    amp_s = math.exp(-alpha * (4.0 * dims * math.sin(0.5 * k * h) ** 2 / (h * h)) * t_final)
    amp_d = (1.0 - 4.0 * dims * r_eff * math.sin(0.5 * k * h) ** 2) ** steps
    return {"n": n, "dims": dims, "bc": bc, "h": h, "dt": dt, "r": r_eff, "steps": steps,
            "t_final": t_final,
            "err_exact": rel_err(u, amp_c * shape0),
            "err_semidiscrete": rel_err(u, amp_s * shape0),
            "err_discrete": rel_err(u, amp_d * shape0)}


def slope(x: list[float], y: list[float]) -> float:
    """Least-squares slope of log y against log x (``validation.py:31-33``)."""
    return float(np.polyfit(np.log(x), np.log(y), 1)[0])


def spatial_convergence(dims: int, ns: list[int], r: float = 0.1, t_final: float = 0.02,
                        bc: str = "dirichlet") -> dict[str, Any]:
    runs = [fourier_run(n, dims, bc, r, t_final) for n in ns]
    return {"runs": runs, "order": slope([x["h"] for x in runs], [x["err_exact"] for x in runs])}


def temporal_convergence(dims: int, n: int, steps_list: list[int], t_final: float = 0.02,
                         bc: str = "dirichlet") -> dict[str, Any]:
    h = 1.0 / (n - 1) if bc == "dirichlet" else 1.0 / n
    runs = [fourier_run(n, dims, bc, t_final / s / (h * h), t_final) for s in steps_list]
    return {"runs": runs,
            "order": slope([x["dt"] for x in runs], [x["err_semidiscrete"] for x in runs])}


def jacobi_iterations(n: int, tol: float) -> int:
    """``ceil(ln tol / ln cos(pi h))`` (``validation.py:89-93``)."""
    rho = math.cos(math.pi / (n - 1))
    return int(math.ceil(math.log(tol) / math.log(rho)))


def solve_laplace(exact: np.ndarray, tol: float, source: np.ndarray | None = None
                  ) -> tuple[np.ndarray, int]:
    """Fixed-count Jacobi from a zero interior, even sweep count (``validation.py:96-110``)."""
    d = exact.ndim
    u0 = exact.astype(np.float64).copy()
    u0[(slice(1, -1),) * d] = 0.0
    it = jacobi_iterations(exact.shape[0], tol)
    it += it % 2
    st = NumpyStepper(u0, 0.0, 0.5 / d, "dirichlet", source)
    st.step(it)
    return st.get(), it


def harmonic_sinh(n: int, dims: int) -> np.ndarray:
    a = np.arange(n, dtype=np.float64) / (n - 1)
    pts = np.meshgrid(*([a] * dims), indexing="ij")
    if dims == 2:
        y, x = pts
        return np.sin(math.pi * x) * np.sinh(math.pi * y) / math.sinh(math.pi)
    z, y, x = pts
    q = math.sqrt(2.0) * math.pi
    return np.sin(math.pi * x) * np.sin(math.pi * y) * np.sinh(q * z) / math.sinh(q)


def harmonic_quadratic(n: int, dims: int) -> np.ndarray:
    a = np.arange(n, dtype=np.float64) / (n - 1)
    pts = np.meshgrid(*([a] * dims), indexing="ij")
    if dims == 2:
        y, x = pts
        return x * x - y * y
    z, y, x = pts
    return x * x + y * y - 2.0 * z * z


def laplace_convergence(dims: int, ns: list[int], tol: float = 1e-10) -> dict[str, Any]:
    runs = []
    for n in ns:
        ex = harmonic_sinh(n, dims)
        u, it = solve_laplace(ex, tol)
        runs.append({"n": n, "h": 1.0 / (n - 1), "iterations": it, "err": rel_err(u, ex)})
    return {"runs": runs, "order": slope([x["h"] for x in runs], [x["err"] for x in runs])}


def laplace_quadratic(dims: int, n: int, tol: float = 1e-12) -> dict[str, Any]:
    ex = harmonic_quadratic(n, dims)
    u, it = solve_laplace(ex, tol)
    return {"n": n, "iterations": it, "err": rel_err(u, ex)}


def poisson_convergence(dims: int, ns: list[int], tol: float = 1e-10) -> dict[str, Any]:
    """Jacobi for ``lap u = f``, ``b = -h^2 f / (2d)`` (``validation.py:135-147``)."""
    runs = []
    for n in ns:
        ex = sine_mode(n, dims, "dirichlet")
        f = -dims * math.pi ** 2 * ex
        h = 1.0 / (n - 1)
        u, it = solve_laplace(ex, tol, source=-h * h * f / (2 * dims))
        runs.append({"n": n, "h": h, "iterations": it, "err": rel_err(u, ex)})
    return {"runs": runs, "order": slope([x["h"] for x in runs], [x["err"] for x in runs])}
