"""The FP64 solutions batched and served results are measured against (CPU only)."""

import numpy as np
import pytest

from oracle import numpy_ref as nr
from pyheat3d_amd import fp64
from pyheat3d_amd.analytic import splitmix_field


@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
@pytest.mark.parametrize("shape", [(20, 21, 22), (31, 40)])
def test_heat_equals_the_oracle_in_float64(bc, shape):
    u0 = splitmix_field(shape, 4, 0)
    c0, c1 = (0.1, 0.15) if len(shape) == 3 else (0.2, 0.2)
    ref = u0.astype(np.float64)
    for _ in range(7):
        ref = nr.reference_step(ref, float(np.float32(c0)), float(np.float32(c1)), bc)
    assert np.array_equal(fp64.heat(u0, c0, c1, bc, 7), ref)


@pytest.mark.parametrize("shape", [(30, 33), (14, 15, 16), (9, 20, 7)])
@pytest.mark.parametrize("with_source", [False, True])
def test_dirichlet_solution_is_the_jacobi_fixed_point(shape, with_source):
    """The DST solve satisfies u = c1 S(u) + b to round-off (the oracle's FP64 residual), and
    float64 Jacobi sweeps from the boundary data converge onto it."""
    d = len(shape)
    c1 = float(np.float32(1 / (2 * d)))
    g = splitmix_field(shape, 3, 1).astype(np.float64)
    b = np.random.default_rng(1).random(shape) * 1e-3 if with_source else None
    u = fp64.dirichlet_solution(g, c1, b)
    core = (slice(1, -1),) * d
    assert np.array_equal(u[~np.pad(np.ones([n - 2 for n in shape], bool), 1)],
                          g[~np.pad(np.ones([n - 2 for n in shape], bool), 1)])  # walls kept
    r0 = nr.residual_fp64(g, c1, "dirichlet", b)
    assert nr.residual_fp64(u, c1, "dirichlet", b) <= 1e-26 * max(r0, 1.0)
    v = g.copy()
    for _ in range(8000):
        v = nr.reference_step(v, 0.0, c1, "dirichlet", b)
    assert fp64.rel_err(v, u) < 1e-12
    assert np.abs(u[core]).max() > 0


def test_rel_err_is_the_max_norm_deviation():
    ref = np.array([1.0, -4.0, 2.0])
    assert fp64.rel_err(ref + np.array([0.0, 0.5, 0.0]), ref) == pytest.approx(0.125)
