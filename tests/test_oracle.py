"""Pins the CPU oracle (oracle/) before anything is compared against it. CPU only.

* NumPy restatement == scalar C restatement, bit for bit, on ragged shapes (both BCs, fp32 and
  fp64, with and without a source); a planted reordering of the sum is detected.
* Closed forms and invariants the reference's own tests use (tests/test_reference.py:13-89).
* The FP64 validation numbers the reference committed (results/validation.json, extracted to
  tests/golden/validation_golden.json).
* Storage-codec known answers and the rounding cadence of the emulated reference.
* The FP64 residual against its direct definition and the Jacobi-eigenmode closed form.
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import cref
from oracle import numpy_ref as nr

GOLD = json.loads((Path(__file__).parent / "golden" / "validation_golden.json").read_text())
SHAPES = [(11, 13, 9), (5, 6, 7), (3, 3, 3), (9, 17), (5, 7), (3, 4)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("with_source", [False, True])
def test_numpy_equals_scalar_c_bitwise(shape, bc, dtype, with_source):
    rng = np.random.default_rng(1)
    u = rng.random(shape).astype(dtype)
    b = rng.random(shape).astype(dtype) * dtype(1e-3) if with_source else None
    a = c = u
    for _ in range(5):
        a = nr.reference_step(a, 0.4, 0.1, bc, b)
        c = cref.step(c, 0.4, 0.1, bc, b)
    assert np.array_equal(a, c)
    st = nr.NumpyStepper(u, 0.4, 0.1, bc, b)
    st.step(5)
    assert np.array_equal(st.get(), a)


def test_planted_reordering_is_detected():
    """Regrouping the sum changes float32 bits: the C comparison pins the bracketing."""
    rng = np.random.default_rng(5)
    u = rng.random((40, 50), dtype=np.float32)
    good = cref.step(u, 0.2, 0.2, "periodic")
    pairs = nr.shift_pairs(u, True)
    (xm, xp), (ym, yp) = pairs
    s = (xm + ym) + (xp + yp)  # same terms, wrong pairing
    bad = np.float32(0.2) * u + np.float32(0.2) * s
    assert not np.array_equal(good, bad)
    u3 = rng.random((8, 9, 10), dtype=np.float32)
    good3 = cref.step(u3, 0.1, 0.15, "periodic")
    (xm, xp), (ym, yp), (zm, zp) = nr.shift_pairs(u3, True)
    s3 = (xm + xp) + ((ym + yp) + (zm + zp))  # right pairs, wrong association
    assert not np.array_equal(good3, np.float32(0.1) * u3 + np.float32(0.15) * s3)
    fused = (np.float32(0.2) * u.astype(np.float64) + np.float32(0.2) * (
        (pairs[0][0] + pairs[0][1]) + (pairs[1][0] + pairs[1][1])).astype(np.float64))
    assert not np.array_equal(good, fused.astype(np.float32))  # one rounding instead of three


def test_planted_dropped_term_is_detected():
    rng = np.random.default_rng(6)
    u = rng.random((7, 8, 9), dtype=np.float32)
    good = cref.step(u, 0.1, 0.15, "dirichlet")
    pairs = nr.shift_pairs(u, False)
    s = (pairs[0][0] + pairs[0][1]) + (pairs[1][0] + pairs[1][1])  # z pair dropped
    bad = u.copy()
    bad[1:-1, 1:-1, 1:-1] = np.float32(0.1) * u[1:-1, 1:-1, 1:-1] + np.float32(0.15) * s
    assert not np.array_equal(good, bad)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
def test_fully_discrete_mode_is_exact(dims, bc):
    n = 16 if bc == "periodic" else 17
    run = nr.fourier_run(n, dims, bc, 0.1, 0.01)
    assert run["err_discrete"] < 1e-12


def test_dirichlet_boundary_untouched():
    rng = np.random.default_rng(2)
    u0 = rng.random((10, 12, 14))
    st = nr.NumpyStepper(u0, 1 - 6 * 0.15, 0.15, "dirichlet")
    st.step(7)
    u = st.get()
    for ax in range(3):
        for idx in (0, -1):
            np.testing.assert_array_equal(np.take(u, idx, ax), np.take(u0, idx, ax))


def test_periodic_heat_conserves_mean():
    rng = np.random.default_rng(3)
    u0 = rng.random((32, 24))
    st = nr.NumpyStepper(u0, 1 - 4 * 0.2, 0.2, "periodic")
    st.step(50)
    assert st.get().mean() == pytest.approx(u0.mean(), rel=1e-13)


def test_maximum_principle():
    rng = np.random.default_rng(4)
    u0 = rng.random((20, 20, 20))
    st = nr.NumpyStepper(u0, 0.0, 1 / 6, "dirichlet")
    st.step(20)
    u = st.get()
    assert u.max() <= u0.max() + 1e-15 and u.min() >= u0.min() - 1e-15


# -- the reference's committed FP64 validation numbers ------------------------------------------

def _close(a, b, rtol=1e-9):
    return abs(a - b) <= rtol * abs(b)


@pytest.mark.parametrize("study", range(2))  # 2D and 3D Dirichlet spatial studies
def test_spatial_study_matches_reference_numbers(study):
    g = GOLD["spatial"][study]
    for run in g["runs"]:
        if run["n"] ** g["dims"] * run["steps"] > 3e8:
            continue  # the largest grids are left to the slow tier
        r = nr.fourier_run(run["n"], g["dims"], g["bc"], g["r"], g["t_final"])
        assert r["steps"] == run["steps"]
        assert _close(r["err_exact"], run["err_exact"])
        assert _close(r["err_semidiscrete"], run["err_semidiscrete"])
        assert r["err_discrete"] < 1e-12 and run["err_discrete"] < 1e-12


def test_periodic_spatial_study_matches_reference_numbers():
    g = GOLD["spatial"][2]
    for run in g["runs"][:3]:
        r = nr.fourier_run(run["n"], g["dims"], g["bc"], g["r"], g["t_final"])
        assert r["steps"] == run["steps"]
        assert _close(r["err_exact"], run["err_exact"])


def test_temporal_study_matches_reference_numbers():
    g = GOLD["temporal"][1]  # 3D, n = 33
    h = 1.0 / (g["n"] - 1)
    for run in g["runs"][:3]:
        r = nr.fourier_run(g["n"], g["dims"], g["bc"], g["t_final"] / run["steps"] / (h * h),
                           g["t_final"])
        assert r["steps"] == run["steps"]
        assert _close(r["err_semidiscrete"], run["err_semidiscrete"])


def test_laplace_and_poisson_match_reference_numbers():
    for g in GOLD["laplace"]:
        for run in g["runs"][:3]:
            ex = nr.harmonic_sinh(run["n"], g["dims"])
            u, it = nr.solve_laplace(ex, 1e-10)
            assert it == run["iterations"]
            assert _close(nr.rel_err(u, ex), run["err"])
    for g in GOLD["poisson"]:
        for run in g["runs"][:3]:
            res = nr.poisson_convergence(g["dims"], [run["n"]])["runs"][0]
            assert res["iterations"] == run["iterations"]
            assert _close(res["err"], run["err"])


def test_laplace_quadratic_matches_reference_numbers():
    q = GOLD["laplace_quadratic"][0]
    r = nr.laplace_quadratic(q["dims"], q["n"])
    assert r["iterations"] == q["iterations"]
    assert r["err"] < 1e-13 and q["err"] < 1e-13


# -- storage codec --------------------------------------------------------------------------

def _f(bits):
    return np.array(bits, dtype=np.uint32).view(np.float32)


def test_bf16_known_answers():
    x = _f([0x3F800000, 0x3F808000, 0x3F818000, 0x3F808001, 0x7F7FFFFF, 0x00000001, 0xBF800000])
    want = np.array([0x3F80, 0x3F80, 0x3F82, 0x3F81, 0x7F80, 0x0000, 0xBF80], np.uint16)
    assert np.array_equal(nr.bf16_bits(x), want)
    assert np.array_equal(cref.round_storage(x, "bfloat16")[0], want)


def test_fp16_known_answers():
    vals = np.array([1.0, 65504.0, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                     1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, -2.0, 2.0 ** -14, 1e-8], np.float32)
    want = np.array([0x3C00, 0x7BFF, 0x7C00, 0x0001, 0x0000, 0x0001, 0x3C00, 0x3C02, 0xC000,
                     0x0400, 0x0000], np.uint16)
    assert np.array_equal(vals.astype(np.float16).view(np.uint16), want)
    assert np.array_equal(cref.round_storage(vals, "float16")[0], want)


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_codec_numpy_equals_c_on_random_bit_patterns(storage):
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    _, back = cref.round_storage(x, storage)
    assert np.array_equal(back, nr.round_storage(x, storage), equal_nan=True)


def test_emulated_reference_cadence():
    rng = np.random.default_rng(12)
    u0 = rng.random((12, 14), dtype=np.float32)
    per_step = np.asarray(u0)
    per_step = nr.round_storage(per_step, "bfloat16")
    for _ in range(8):
        per_step = nr.round_storage(nr.reference_step(per_step, 0.2, 0.2, "periodic"), "bfloat16")
    assert np.array_equal(nr.emulated_reference(u0, 0.2, 0.2, "periodic", 8, "bfloat16", 1),
                          per_step)
    k4 = nr.emulated_reference(u0, 0.2, 0.2, "periodic", 8, "bfloat16", 4)
    assert not np.array_equal(k4, per_step)  # fewer roundings with fused steps
    f32 = nr.emulated_reference(u0, 0.2, 0.2, "periodic", 8)
    assert np.array_equal(f32, cref.run_f32(u0, 0.2, 0.2, "periodic", 8))


@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
def test_window_reference_equals_the_full_grid_oracle(bc):
    """Light-cone windows (corners, faces, centre, across the periodic wrap) equal the same
    sites of the full-grid oracle bit for bit; with one layer of margin too few they do not."""
    rng = np.random.default_rng(23)
    steps, size = 5, 6
    for shape in ((30, 31, 29), (33, 40)):
        u0 = rng.random(shape, dtype=np.float32)
        c0, c1 = 0.1, 0.9 / (2 * len(shape))
        full = nr.emulated_reference(u0, c0, c1, bc, steps)
        ends = [n - size for n in shape]
        lows = [(0,) * len(shape), tuple(ends), tuple(n // 2 for n in ends),
                tuple(0 if i % 2 else e for i, e in enumerate(ends))]
        if bc == "periodic":
            lows += [tuple(n - size // 2 for n in shape), tuple(n - 1 for n in shape)]
        for lo in lows:
            ref, idx = nr.window_reference(u0, lo, size, c0, c1, bc, steps)
            assert ref.shape == (size,) * len(shape)
            assert np.array_equal(ref, full[np.ix_(*idx)]), (shape, lo)
        # a margin one layer short leaves stale values on the faces that are not walls
        lo = tuple(n // 2 for n in ends)
        short, idx = nr.window_reference(u0, lo, size, c0, c1, bc, steps,
                                         margin=steps - 1)
        assert not np.array_equal(short, full[np.ix_(*idx)]), shape
        if bc == "dirichlet":
            with pytest.raises(ValueError):
                nr.window_reference(u0, (shape[0] - 1,) + (0,) * (len(shape) - 1), size, c0, c1,
                                    bc, steps)


def test_window_reference_with_storage_rounding_and_source():
    """Windows of the bf16 / fp16 storage oracle with fused passes and a source term (the
    Jacobi Poisson setting) equal the full-grid oracle's sites bit for bit."""
    rng = np.random.default_rng(24)
    shape, size = (26, 27, 28), 6
    u0 = rng.random(shape, dtype=np.float32)
    b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
    for storage, k, steps in (("bfloat16", 4, 7), ("float16", 1, 6), ("float16", 4, 8)):
        full = nr.emulated_reference(u0, 0.0, 1 / 6, "dirichlet", steps, storage, k, b)
        for lo in ((0, 0, 0), (10, 11, 12), (20, 0, 22)):
            ref, idx = nr.window_reference(u0, lo, size, 0.0, 1 / 6, "dirichlet", steps,
                                           storage, k, b)
            assert np.array_equal(ref, full[np.ix_(*idx)]), (storage, k, lo)


@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
def test_window_reference_c_engine_equals_numpy(bc):
    """float32 windows stepped by the scalar C loop equal the NumPy windows and the full-grid
    oracle bit for bit, with a source and with the light cone clipped at the walls or wrapped."""
    rng = np.random.default_rng(25)
    shape, size, steps = (22, 23, 24), 5, 6
    u0 = rng.random(shape, dtype=np.float32)
    b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
    full = nr.emulated_reference(u0, 0.1, 0.15, bc, steps, b=b)
    for lo in ((0, 0, 0), (9, 10, 11), (17, 0, 19)):
        a, idx = nr.window_reference(u0, lo, size, 0.1, 0.15, bc, steps, b=b, engine="c")
        c, _ = nr.window_reference(u0, lo, size, 0.1, 0.15, bc, steps, b=b)
        assert np.array_equal(a, c) and np.array_equal(a, full[np.ix_(*idx)]), lo


# -- FP64 residual ----------------------------------------------------------------------------

@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
@pytest.mark.parametrize("shape", [(9, 10, 11), (13, 17)])
def test_residual_numpy_equals_c(bc, shape):
    rng = np.random.default_rng(13)
    u = rng.random(shape, dtype=np.float32)
    b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
    d = len(shape)
    for bb in (None, b):
        r1 = nr.residual_fp64(u, 1 / (2 * d), bc, bb)
        r2 = cref.residual_fp64(u, 1 / (2 * d), bc, bb)
        assert r1 == pytest.approx(r2, rel=1e-12)


def test_residual_brute_force_tiny():
    """Pure-Python loop over every site of a 4x5x6 periodic grid (wrapped indices by %)."""
    rng = np.random.default_rng(14)
    u = rng.random((4, 5, 6), dtype=np.float32)
    c1 = float(np.float32(1 / 6))
    acc = 0.0
    for k in range(4):
        for j in range(5):
            for i in range(6):
                s = (float(u[k, j, (i - 1) % 6]) + float(u[k, j, (i + 1) % 6]) +
                     float(u[k, (j - 1) % 5, i]) + float(u[k, (j + 1) % 5, i]) +
                     float(u[(k - 1) % 4, j, i]) + float(u[(k + 1) % 4, j, i]))
                r = c1 * s - float(u[k, j, i])
                acc += r * r
    assert nr.residual_fp64(u, 1 / 6, "periodic") == pytest.approx(acc, rel=1e-12)


def test_residual_jacobi_eigenmode_closed_form():
    """From zero, Jacobi on the sine eigenmode has relative residual exactly mu^s.

    With the float32 coefficient c1 = f32(1/6) that the kernels use, the Jacobi operator maps
    the mode v to mu v with mu = 6 c1 cos(pi h), so u_s = b (1 + mu + ... + mu^(s-1)) and the
    residual b + c1 S(u_s) - u_s equals b mu^s.
    """
    n, d = 17, 3
    h = 1.0 / (n - 1)
    c1 = float(np.float32(1 / 6))
    mu = 2 * d * c1 * math.cos(math.pi * h)
    v = nr.sine_mode(n, d, "dirichlet")
    b = (h * h * math.pi ** 2 / 2) * v
    st = nr.NumpyStepper(np.zeros_like(v), 0.0, c1, "dirichlet", b)
    r0 = nr.residual_fp64(st.get(), c1, "dirichlet", b)
    for s in (10, 40):
        st.step(s - st.steps_done)
        ratio = math.sqrt(nr.residual_fp64(st.get(), c1, "dirichlet", b) / r0)
        assert ratio == pytest.approx(mu ** s, rel=1e-9)
