"""Validation studies of ``hip-cdna4`` variants against closed forms (no CPU oracle involved).

Restates the reference's ``src/pyheat3d/validation.py`` studies for float32 GPU variants:

* :func:`fourier_run` -- the decaying sine mode vs the continuous, semi-discrete and fully
  discrete (``g**n``) amplitudes (``validation.py:36-64``);
* :func:`spatial_convergence` / :func:`temporal_convergence` -- observed orders 2 and 1
  (``validation.py:67-86``); in float32 the finest grids approach round-off, so the studies
  use grids where the discretisation error dominates;
* :func:`laplace_convergence` / :func:`poisson_convergence` -- Jacobi for the reference's
  fixed sweep count (``jacobi_iterations``) on a smooth harmonic function and on a manufactured
  Poisson problem: observed order 2 (``validation.py:89-147``);
* :func:`laplace_quadratic` -- Jacobi to tolerance on a quadratic harmonic, which the 5/7-point
  stencil reproduces exactly (``validation.py:126-132``);
* :func:`poisson_eigen` -- Jacobi on ``-lap u = d pi^2 prod sin(pi x_i)``, whose discrete
  solution is ``alpha * v`` in closed form (the FP64 solution), reporting the error of each
  storage type.

Cross-checks against the NumPy oracle live in the test suite, not here.
"""

from __future__ import annotations

import math
from typing import Any

import numpy as np

from . import analytic as an
from .jacobi import solve
from .problem import heat_stencil
from .registry import VARIANTS, availability, make_stepper


def _rel(u: np.ndarray, ref: np.ndarray) -> float:
    return float(np.max(np.abs(u - ref)) / max(np.max(np.abs(ref)), 1e-300))


def slope(x: list[float], y: list[float]) -> float:
    return float(np.polyfit(np.log(x), np.log(y), 1)[0])


def fourier_run(key: str, n: int, dims: int, bc: str, r: float, t_final: float,
                alpha: float = 1.0) -> dict[str, Any]:
    h = an.grid_spacing(n, bc)
    steps = max(1, int(math.ceil(t_final / (r * h * h / alpha))))
    dt = t_final / steps
    r_eff = alpha * dt / (h * h)
    shape0 = an.fourier_mode(n, dims, bc)
    st = make_stepper(key, shape0.astype(np.float32), heat_stencil(dims, r_eff, bc))
    try:
        st.step(steps)
        u = st.get().astype(np.float64)
    finally:
        st.close()
    return {"variant": key, "n": n, "dims": dims, "bc": bc, "h": h, "dt": dt, "r": r_eff,
            "steps": steps, "t_final": t_final,
            "err_exact": _rel(u, an.decay_continuous(t_final, dims, bc, alpha=alpha) * shape0),
            "err_semidiscrete": _rel(u, an.decay_semidiscrete(t_final, n, dims, bc,
                                                              alpha=alpha) * shape0),
            "err_discrete": _rel(u, an.ftcs_amplification(r_eff, n, dims, bc) ** steps * shape0)}


def spatial_convergence(key: str, dims: int, ns: list[int], r: float = 0.1,
                        t_final: float = 0.02, bc: str = "dirichlet") -> dict[str, Any]:
    runs = [fourier_run(key, n, dims, bc, r, t_final) for n in ns]
    return {"dims": dims, "bc": bc, "variant": key, "runs": runs,
            "order": slope([x["h"] for x in runs], [x["err_exact"] for x in runs])}


def temporal_convergence(key: str, dims: int, n: int, steps_list: list[int],
                         t_final: float = 0.02, bc: str = "dirichlet") -> dict[str, Any]:
    h = an.grid_spacing(n, bc)
    runs = [fourier_run(key, n, dims, bc, t_final / s / (h * h), t_final) for s in steps_list]
    return {"dims": dims, "bc": bc, "variant": key, "n": n, "runs": runs,
            "order": slope([x["dt"] for x in runs], [x["err_semidiscrete"] for x in runs])}


def jacobi_iterations(n: int, tol: float) -> int:
    """Sweeps for Jacobi (spectral radius cos(pi h)) to damp an O(1) error to ``tol``
    (``validation.py:89-93``)."""
    return int(math.ceil(math.log(tol) / math.log(math.cos(math.pi / (n - 1)))))


def solve_fixed(key: str, exact: np.ndarray, tol: float,
                source: np.ndarray | None = None) -> tuple[np.ndarray, int]:
    """Jacobi from a zero interior with the boundary values of ``exact`` for the reference's
    fixed (even) sweep count (``validation.py:96-110``), in the variant's float32 arithmetic."""
    from .problem import jacobi_stencil

    dims = exact.ndim
    u0 = exact.astype(np.float32)
    u0[(slice(1, -1),) * dims] = 0.0
    sweeps = jacobi_iterations(exact.shape[0], tol)
    sweeps += sweeps % 2
    b = None if source is None else np.asarray(source, dtype=np.float32)
    st = make_stepper(key, u0, jacobi_stencil(dims), b)
    try:
        st.step(sweeps)
        u = st.get().astype(np.float64)
    finally:
        st.close()
    return u, sweeps


def _jacobi_orders(runs: list[dict[str, Any]]) -> dict[str, float]:
    """Observed order over the grids where the discretisation error dominates float32.

    Jacobi converges to the fixed point of its rounded map, at most eps32 / (1 - cos(pi h)) from
    the exact one (the per-sweep rounding times the iteration's sensitivity; each run records it
    as ``fp32_bound``); a grid where that bound reaches half the error has left the asymptotic
    range in float32 (2D n = 129: 2e-4 against a 1.7e-5 error). ``order`` fits the other grids
    (all grids if fewer than three remain), ``order_all_grids`` every grid.
    """
    for r in runs:
        r["fp32_bound"] = 2.0 ** -24 / (1.0 - math.cos(math.pi / (r["n"] - 1)))
    dom = [r for r in runs if r["fp32_bound"] < 0.5 * r["err"]]
    if len(dom) < 3:
        dom = runs
    return {"order": slope([x["h"] for x in dom], [x["err"] for x in dom]),
            "order_grids": [x["n"] for x in dom],
            "order_all_grids": slope([x["h"] for x in runs], [x["err"] for x in runs])}


def laplace_convergence(key: str, dims: int, ns: list[int], tol: float = 1e-10) -> dict[str, Any]:
    """Converged Jacobi solution vs the smooth harmonic function (``analytic.harmonic_sinh``):
    order 2 in h (``validation.py:113-123``)."""
    runs = []
    for n in ns:
        ex = an.harmonic_sinh(n, dims)
        u, it = solve_fixed(key, ex, tol)
        runs.append({"n": n, "h": 1.0 / (n - 1), "iterations": it, "err": _rel(u, ex)})
    return {"dims": dims, "variant": key, "tol": tol, "runs": runs, **_jacobi_orders(runs)}


def poisson_convergence(key: str, dims: int, ns: list[int], tol: float = 1e-10) -> dict[str, Any]:
    """Jacobi for ``lap u = f`` with the manufactured solution of ``analytic.poisson_manufactured``
    and source ``b = -h^2 f / (2d)`` (rounded to float32): order 2 in h
    (``validation.py:135-147``)."""
    runs = []
    for n in ns:
        ex, f = an.poisson_manufactured(n, dims)
        h = 1.0 / (n - 1)
        u, it = solve_fixed(key, ex, tol, source=-h * h * f / (2 * dims))
        runs.append({"n": n, "h": h, "iterations": it, "err": _rel(u, ex)})
    return {"dims": dims, "variant": key, "tol": tol, "runs": runs, **_jacobi_orders(runs)}


def laplace_quadratic(key: str, dims: int, n: int, tol: float = 1e-6) -> dict[str, Any]:
    ex = an.harmonic_quadratic(n, dims)
    u0 = ex.astype(np.float32)
    u0[(slice(1, -1),) * dims] = 0.0
    res = solve(key, u0, tol=tol, check_every=64)
    return {"variant": key, "dims": dims, "n": n, "status": res.status, "sweeps": res.sweeps,
            "rel_residual": res.rel_residual, "err": _rel(res.u.astype(np.float64), ex)}


def poisson_eigen(key: str, n: int, dims: int = 3, tol: float = 1e-6,
                  check_every: int = 100, max_sweeps: int = 10 ** 7,
                  progress=None) -> dict[str, Any]:
    """Jacobi from zero on the manufactured eigen-problem; error vs the FP64 solution alpha*v.

    The relative residual of the exact iteration is ``mu^s`` with ``mu = 2 d c1 cos(pi h)``
    (``c1`` the float32 coefficient), reported as ``exact_rel_residual`` beside the measured one
    so the distance to the storage floor is visible.
    """
    v, b, alpha = an.poisson_eigen_source(n, dims)
    u0 = np.zeros(v.shape, np.float32)
    b32 = b.astype(np.float32)
    del b
    res = solve(key, u0, b32, tol=tol, check_every=check_every, max_sweeps=max_sweeps,
                progress=progress)
    c1 = float(np.float32(0.5 / dims))
    mu = 2 * dims * c1 * math.cos(math.pi / (n - 1))
    err_fp64 = _rel(res.u.astype(np.float64), alpha * v)
    err_pde = _rel(res.u.astype(np.float64), v)
    return {"variant": key, "storage": VARIANTS[key].storage, "n": n, "dims": dims,
            "status": res.status, "sweeps": res.sweeps, "rel_residual": res.rel_residual,
            "exact_rel_residual": mu ** res.sweeps, "seconds": res.seconds,
            "predicted_sweeps": math.ceil(math.log(tol) / math.log(mu)),
            "err_vs_fp64": err_fp64, "err_vs_pde": err_pde,
            "history": res.history[:: max(1, len(res.history) // 64)]}


def run_all(quick: bool = True) -> dict[str, Any]:
    """Every study on the available float32 ``hip-cdna4`` variants."""
    out: dict[str, Any] = {"quick": quick}
    keys2 = [k for k in ("hip-cdna4/naive", "hip-cdna4/ymarch") if availability(k)[0]]
    keys3 = [k for k in ("hip-cdna4/naive", "hip-cdna4/zmarch") if availability(k)[0]]
    if not keys3:
        raise RuntimeError("no hip-cdna4 variant is available on this machine")
    ns2 = [17, 33, 65] if quick else [17, 33, 65, 129]
    ns3 = [9, 17, 33]
    out["spatial"] = [spatial_convergence(k, 2, ns2) for k in keys2] + \
                     [spatial_convergence(k, 3, ns3) for k in keys3]
    out["temporal"] = [temporal_convergence(k, 2, 17, [64, 128, 256]) for k in keys2[:1]]
    # the reference's grids (validation.py:258-262); quick: the 2D studies on three grids
    out["laplace"] = [laplace_convergence(k, 2, [9, 17, 33] if quick else [9, 17, 33, 65, 129])
                      for k in keys2[:1]]
    out["poisson"] = [poisson_convergence(k, 2, [9, 17, 33] if quick else [9, 17, 33, 65])
                      for k in keys2[:1]]
    if not quick:
        out["laplace"] += [laplace_convergence(k, 3, [9, 17, 33, 65]) for k in keys3[:1]]
        out["poisson"] += [poisson_convergence(k, 3, [9, 17, 33]) for k in keys3[:1]]
    out["laplace_quadratic"] = [laplace_quadratic(k, 2, 33) for k in keys2[:1]] + \
                               [laplace_quadratic(k, 3, 17) for k in keys3[:1]]
    out["poisson_eigen"] = [poisson_eigen(k, 33) for k in keys3]
    return out
