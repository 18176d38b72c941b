"""Problem definition, analytic fields, traffic model, workload generators, CLI. CPU only.

Mirrors the reference's tests/test_problem.py and tests/test_cli.py where they apply.
"""

import json
import math

import numpy as np
import pytest

from pyheat3d_amd import analytic as an
from pyheat3d_amd import traffic
from pyheat3d_amd.batch import Request, make_mix
from pyheat3d_amd.jacobi import jacobi_rate
from pyheat3d_amd.problem import heat_stencil, jacobi_stencil, stability_limit, Stencil


def test_stability_limit():
    assert stability_limit(2) == 0.25
    assert stability_limit(3) == pytest.approx(1 / 6)


def test_heat_coefficients_and_limit():
    st = heat_stencil(3, 0.1)
    assert st.c0 == pytest.approx(0.4) and st.c1 == 0.1
    heat_stencil(3, 1 / 6)
    with pytest.raises(ValueError):
        heat_stencil(3, 0.17)
    with pytest.raises(ValueError):
        heat_stencil(2, 0.26, "periodic")
    with pytest.raises(ValueError):
        heat_stencil(3, 0.1, "neumann")
    with pytest.raises(ValueError):
        Stencil(0.0, 0.25, "open")


def test_jacobi_coefficients():
    assert jacobi_stencil(2).c1 == 0.25
    assert jacobi_stencil(3).c0 == 0.0
    assert jacobi_stencil(3).c1 == 1 / 6


def test_workload_coefficients():
    # config 1/2: r = 0.15 in 3D; config 4: r = 0.2 in 2D
    st3 = heat_stencil(3, 0.15)
    assert st3.c0 == pytest.approx(0.1) and st3.c1 == 0.15
    st2 = heat_stencil(2, 0.2)
    assert st2.c0 == pytest.approx(0.2) and st2.c1 == 0.2


def test_fourier_mode_is_zero_on_dirichlet_boundary():
    u = an.fourier_mode(17, 3, "dirichlet")
    assert u.shape == (17, 17, 17)
    assert np.abs(u[0]).max() == 0 and np.abs(u[:, :, -1]).max() == 0
    assert u.max() == pytest.approx(1.0)


def test_decay_rates_agree_for_small_h():
    t = 0.01
    c = an.decay_continuous(t, 3, "dirichlet")
    s = an.decay_semidiscrete(t, 1025, 3, "dirichlet")
    assert s == pytest.approx(c, rel=1e-5)
    g = an.ftcs_amplification(0.1, 1025, 3, "dirichlet")
    steps = round(t / (0.1 / 1024 ** 2))
    assert g ** steps == pytest.approx(c, rel=1e-4)


def test_harmonic_functions_are_discretely_harmonic():
    q = an.harmonic_quadratic(9, 3)
    lap = (q[1:-1, 1:-1, :-2] + q[1:-1, 1:-1, 2:] + q[1:-1, :-2, 1:-1] + q[1:-1, 2:, 1:-1]
           + q[:-2, 1:-1, 1:-1] + q[2:, 1:-1, 1:-1] - 6 * q[1:-1, 1:-1, 1:-1])
    assert np.abs(lap).max() < 1e-14


def test_poisson_eigen_source_is_the_discrete_solution():
    """alpha*v solves the 7-point system: alpha*v = (1/6) S(alpha*v) + b on the interior."""
    n = 21
    v, b, alpha = an.poisson_eigen_source(n, 3)
    u = alpha * v
    c = (slice(1, -1),) * 3
    s = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1]
         + u[:-2, 1:-1, 1:-1] + u[2:, 1:-1, 1:-1])
    assert np.abs(s / 6 + b[c] - u[c]).max() < 1e-14
    assert an.jacobi_eigen_iterate(n, 10 ** 6) == pytest.approx(alpha)


def test_sine_modes_init_projects_on_each_mode():
    n = 32
    u = an.sine_modes_periodic(n, 3, noise=0.0)
    x = np.arange(n) / n
    for m in (1, 2, 3):
        s = np.sin(2 * np.pi * m * x)
        amp = np.einsum("zyx,z,y,x->", u.astype(np.float64), s, s, s) / (s @ s) ** 3
        assert amp == pytest.approx(1.0, rel=1e-6)


def test_random_dirichlet_recipe():
    u = an.random_dirichlet((6, 7, 8), 0)
    full = np.random.default_rng(0).random((6, 7, 8), dtype=np.float32)
    assert np.array_equal(u[1:-1, 1:-1, 1:-1], full[1:-1, 1:-1, 1:-1])
    assert u[0].max() == 0 and u[:, -1].max() == 0


def test_splitmix_field_kinds():
    a = an.splitmix_field((5, 6, 7), 11, 2)
    assert a.dtype == np.float32 and 0 <= a.min() and a.max() < 1
    assert np.array_equal(a, an.splitmix_field((5, 6, 7), 11, 2))
    k0 = an.splitmix_field((5, 6, 7), 11, 0)
    k1 = an.splitmix_field((5, 6, 7), 11, 1)
    assert np.array_equal(k0 + k1, a)  # complementary masks of the same stream
    assert k0[0].max() == 0 and k1[2, 3, 3] == 0
    # values are multiples of 2^-24 (exact in float32)
    assert np.all((a * 2 ** 24) == np.floor(a * 2 ** 24))


def test_traffic_model():
    assert traffic.minimal_bytes(4) == 8
    assert traffic.minimal_bytes(2) == 4
    assert traffic.minimal_bytes(4, 4) == 2
    assert traffic.minimal_bytes(4, 1, 4) == 12  # Jacobi with a float32 source
    assert traffic.flops_per_update(3) == 8
    assert traffic.roofline_mlups(361.0, 8) == pytest.approx(45125.0)
    tb = traffic.tb_model_bytes(3, 4, 4, 32, 64, 128)
    assert traffic.minimal_bytes(4, 4) < tb < traffic.minimal_bytes(4, 1)
    assert math.isclose(traffic.tb_model_bytes(2, 1, 4, 10 ** 6, 1, 10 ** 6), 8, rel_tol=1e-5)


def test_jacobi_rate():
    assert jacobi_rate((33, 33)) == pytest.approx(math.cos(math.pi / 32))
    assert jacobi_rate((9, 17, 33)) == pytest.approx(
        (math.cos(math.pi / 8) + math.cos(math.pi / 16) + math.cos(math.pi / 32)) / 3)
    assert jacobi_rate((8, 8), "periodic") == 1.0


def test_request_mix_is_deterministic_and_in_range():
    a = make_mix(300, seed=5)
    b = make_mix(300, seed=5)
    assert [(r.shape, r.mode, r.seed) for r in a] == [(r.shape, r.mode, r.seed) for r in b]
    d2 = [r for r in a if r.dims == 2]
    d3 = [r for r in a if r.dims == 3]
    assert 0.5 < len(d2) / len(a) < 0.7
    assert all(64 <= r.shape[0] <= 2048 for r in d2) and all(32 <= r.shape[0] <= 256 for r in d3)
    assert {r.mode for r in a} == {"heat", "jacobi"}
    r = Request(shape=(10, 12), mode="heat")
    assert r.stencil().c1 == 0.2 and r.device_bytes() == 2 * 120 * 4 and r.kind == 0
    assert r.stencil().bc == "dirichlet" and r.work() == 8 * 10 * 200
    p = Request(shape=(10, 12), mode="heat", bc="periodic")
    assert p.stencil().bc == "periodic" and p.work() == 120 * 200
    with pytest.raises(ValueError):
        Request(shape=(10, 12), mode="jacobi", bc="periodic")
    with pytest.raises(ValueError):
        Request(shape=(10, 12), bc="neumann")


def test_cli_list_and_json(capsys):
    from pyheat3d_amd.cli import main
    from pyheat3d_amd.registry import VARIANTS

    assert main(["list"]) == 0
    assert "hip-cdna4/zmarch" in capsys.readouterr().out
    json.dumps({k: v.description for k, v in VARIANTS.items()})


def test_tuning_table_keys_resolve():
    from pyheat3d_amd import tune

    table = tune.load_table()
    assert table, "the committed gfx950 tuning table is missing"
    key = tune.table_key("zmarch", 1, 3, "float32", "periodic", (1024, 1024, 1024))
    assert key in table
    t = tune.lookup("zmarch", 1, 3, "float32", "periodic", (1024, 1024, 1024))
    assert t is not None and t.vec in (1, 2, 4)
    assert tune.table_key("tb", 3, 3, "float32", "dirichlet", (8, 8, 8), source=True) == \
        "tb/k3/3d/float32/dirichlet/src/b9"


def test_tuned_tile_needs_a_lane_vector_that_divides_the_row():
    """An entry serves every shape of its size bucket; for a row its lane vector does not divide
    (an odd width in the same bucket, e.g. 121^3 next to a 128^3 entry) it is not applied and the
    library's defaults pick a valid width, instead of the sweep refusing the tile."""
    from pyheat3d_amd import tune

    checked = 0
    for key, e in tune.load_table().items():
        if key.startswith("batch/") or e["tile"]["vec"] <= 1:   # batch entries: tile ids
            continue
        kind, kk, dims, storage, bc = key.split("/")[:5]
        shape = tuple(e["shape"])
        odd = shape[:-1] + (shape[-1] - 1 if shape[-1] % 2 == 0 else shape[-1],)
        if tune.bucket(odd) != tune.bucket(shape) or odd[-1] % 2 == 0:
            continue
        src = "/src/" in key
        args = (kind, int(kk[1:]), int(dims[0]), storage, bc)
        assert tune.lookup(*args, shape, src) is not None, key
        assert tune.lookup(*args, odd, src) is None, key
        checked += 1
    assert checked > 10


def test_bench_cpu_solve_seconds_runs_the_oracle():
    """The batch CPU baseline's per-solve timer (heat in full, Jacobi extrapolated)."""
    import bench
    from pyheat3d_amd.batch import Request

    for mode, sweeps in (("heat", 30), ("jacobi", 640)):
        q = Request(shape=(12, 14), mode=mode, steps=sweeps, seed=3)
        st = q.stencil()
        t = bench._cpu_solve_seconds((q.shape, q.seed, q.kind, st.c0, st.c1, q.mode, sweeps))
        assert 0.0 < t < 5.0


def test_batch_table_entries_are_valid():
    """Batch entries of the tuning table: a shape id the library builds (0 = its rule; 1-5 in
    2D, 1-6 in 3D) for every size bucket of each fused class, and chunk policies the library
    accepts (read back as ph3_batch_params arguments)."""
    from pyheat3d_amd import tune

    table = tune.load_table()
    tiles = {k: e for k, e in table.items() if k.startswith("batch/tile/")}
    assert tiles and all(0 <= e["tile"] <= max(tune.BATCH_SHAPES[int(k.split("/")[2][0])])
                         for k, e in tiles.items())
    for (d, _k), (lo, hi) in tune.BATCH_CLASSES.items():
        for b in range(tune.bucket((lo,) * d), tune.bucket((hi,) * d) + 1):
            assert f"batch/tile/{d}d/b{b}" in tiles
    pol = tune.batch_policy()
    assert {(d, k) for d, k, _, _ in pol} == set(tune.BATCH_CLASSES)
    assert all(t >= 0 and 0 <= c <= 4096 for _, _, t, c in pol)
    assert tune.batch_tile(3, (256, 256, 256)) == table["batch/tile/3d/b24"]["tile"]
