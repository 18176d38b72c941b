"""Jacobi solves to a residual tolerance on the GPU, with the residual norm in FP64.

The reference has no residual: ``validation.solve_laplace`` runs a fixed sweep count
``ceil(ln tol / ln cos(pi h))`` (``src/pyheat3d/validation.py:89-110``). Here a solve stops on
a measured criterion:

* residual ``r(u) = b + c0*u + c1*S(u) - u`` (the defect of one sweep, proportional to
  ``f - A u`` by ``h^2 / (2d)``), squared 2-norm accumulated in FP64 on the device
  (``ph3_residual``, deterministic order);
* converged when ``||r(u_s)|| <= tol * ||r(u_0)||`` (for a zero initial guess this is
  ``||f - A u|| / ||f||``);
* checked every ``check_every`` sweeps, a fixed property of the solve (never of a batch);
* stagnated when, over ``stall_window`` checks, the residual fell by less than
  ``stall_fraction`` of what the Jacobi rate ``rho = mean_i cos(pi / (n_i - 1))`` predicts:
  low-precision storage stops making progress at a floor (rounding of the field), and in fp16
  or bf16 the increments can round to zero, so the update alone would signal false
  convergence while the true residual stays large;
* capped at ``max_sweeps``.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .problem import jacobi_stencil
from .registry import make_stepper


def jacobi_rate(shape: tuple[int, ...], bc: str = "dirichlet") -> float:
    """Spectral radius of the Jacobi sweep for the slowest mode of a Dirichlet box."""
    if bc != "dirichlet":
        return 1.0
    return sum(math.cos(math.pi / (n - 1)) for n in shape) / len(shape)


@dataclass
class JacobiResult:
    status: str                  # "converged" | "stagnated" | "max_sweeps"
    sweeps: int
    rel_residual: float
    res0_sq: float
    seconds: float
    history: list[tuple[int, float]] = field(default_factory=list)
    u: np.ndarray | None = None


def solve(key: str, u0: np.ndarray, source: np.ndarray | None = None, bc: str = "dirichlet",
          tol: float = 1e-6, check_every: int = 100, max_sweeps: int = 10 ** 7,
          stall_window: int = 20, stall_fraction: float = 0.05, keep: bool = True,
          stepper=None, progress=None) -> JacobiResult:
    """Jacobi sweeps of variant ``key`` from ``u0`` until converged, stagnated or capped."""
    st = stepper or make_stepper(key, u0, jacobi_stencil(u0.ndim, bc), source)
    k = st.steps_per_pass
    check_every = max(k, check_every - check_every % k)
    log_rate = math.log(max(jacobi_rate(st.shape, bc), 1e-300)) * check_every
    t0 = time.perf_counter()
    try:
        r0 = st.residual_sq()
        hist: list[tuple[int, float]] = [(0, 1.0)]
        status, rel = "max_sweeps", 1.0
        if r0 == 0.0:
            status, rel = "converged", 0.0
        sweeps = 0
        while status == "max_sweeps" and sweeps < max_sweeps:
            n = min(check_every, max_sweeps - sweeps)
            st.step(n)
            sweeps += n
            rel = math.sqrt(st.residual_sq() / r0)
            hist.append((sweeps, rel))
            if progress is not None:
                progress(sweeps, rel)
            if rel <= tol:
                status = "converged"
                break
            if len(hist) > stall_window and log_rate < 0:
                old = hist[-1 - stall_window][1]
                drop = math.log(max(rel, 1e-300) / old)  # negative when improving
                if drop > stall_fraction * stall_window * log_rate:
                    status = "stagnated"
                    break
        st.synchronize()
        u = st.get() if keep else None
    finally:
        if stepper is None:
            st.close()
    return JacobiResult(status, sweeps, rel, r0, time.perf_counter() - t0, hist, u)
