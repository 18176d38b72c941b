"""Every ``hip-cdna4`` variant against the CPU oracle on the GPU (reference tests/test_backends.py).

fp32 variants must equal the oracle bit for bit (FMA contraction is off in the library, and the
per-site operation order is the oracle's); the test asserts bitwise equality and, as the
stated fp32 bound, a max relative deviation <= 1e-6. fp16/bf16 variants are compared with the
oracle that rounds to storage after every pass of k steps, within the reference's storage
tolerances (validation.py:153-154).
"""

import numpy as np
import pytest

from oracle import numpy_ref as nr
from pyheat3d_amd import _hip
from pyheat3d_amd.problem import heat_stencil, jacobi_stencil
from pyheat3d_amd.registry import VARIANTS, availability, make_stepper

pytestmark = pytest.mark.gpu

GPU_KEYS = [k for k, v in VARIANTS.items() if v.gpu]
SHAPES = {2: (131, 517), 3: (45, 70, 97)}
STEPS = 2 * 3 * 4 * 8 + 5  # every k of the ladder leaves a remainder


def _need(key, dims):
    if dims not in VARIANTS[key].dims:
        pytest.skip(f"{key} has no {dims}D kernel")
    ok, why = availability(key)
    if not ok:
        pytest.skip(why)


def _run(key, u0, st, steps, source=None):
    print(f"[gpu] {key} shape={u0.shape} bc={st.bc} steps={steps}", flush=True)
    s = make_stepper(key, u0, st, source)
    try:
        s.step(steps)
        s.synchronize()
        return s.get()
    finally:
        s.close()


def _check(key, u, ref):
    var = VARIANTS[key]
    err = nr.rel_err(u, ref)
    bound = 1e-6 if var.storage == "float32" else nr.TOLERANCE[var.storage]
    assert err <= bound, (key, err)
    # same per-site operations and the same round-to-nearest-even storage rounding: every
    # storage type is expected to match its oracle bit for bit
    assert np.array_equal(u, ref), (key, err, int(np.sum(u != ref)))


@pytest.mark.parametrize("bc", ["dirichlet", "periodic"])
@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("key", GPU_KEYS)
def test_matches_oracle(key, dims, bc):
    _need(key, dims)
    rng = np.random.default_rng(7)
    u0 = rng.random(SHAPES[dims], dtype=np.float32)
    st = heat_stencil(dims, 0.9 / (2 * dims), bc)
    var = VARIANTS[key]
    k = var.steps_per_pass if var.storage != "float32" else 1
    ref = nr.emulated_reference(u0, st.c0, st.c1, bc, STEPS, var.storage, k)
    _check(key, _run(key, u0, st, STEPS), ref)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("key", GPU_KEYS)
def test_jacobi_with_source(key, dims):
    _need(key, dims)
    rng = np.random.default_rng(8)
    u0 = rng.random(SHAPES[dims], dtype=np.float32)
    b = (rng.random(SHAPES[dims], dtype=np.float32) * np.float32(1e-3))
    st = jacobi_stencil(dims)
    var = VARIANTS[key]
    k = var.steps_per_pass if var.storage != "float32" else 1
    ref = nr.emulated_reference(u0, st.c0, st.c1, "dirichlet", 24, var.storage, k, b)
    _check(key, _run(key, u0, st, 24, b), ref)


@pytest.mark.parametrize("key", [k for k in GPU_KEYS if 3 in VARIANTS[k].dims
                                 and VARIANTS[k].storage == "float32"])
def test_config1_128cube_dirichlet_100_steps(key):
    """128^3 fp32, walls 0, interior U[0,1) seed 0, r = 0.15, 100 steps: bitwise vs the oracle."""
    _need(key, 3)
    from pyheat3d_amd.analytic import random_dirichlet

    u0 = random_dirichlet((128, 128, 128), 0)
    st = heat_stencil(3, 0.15)
    ref = nr.emulated_reference(u0, st.c0, st.c1, "dirichlet", 100)
    _check(key, _run(key, u0, st, 100), ref)


def test_small_grids_smaller_than_one_tile():
    rng = np.random.default_rng(9)
    for dims, shape in ((2, (5, 7)), (3, (5, 6, 7)), (2, (3, 3)), (3, (3, 4, 3))):
        u0 = rng.random(shape, dtype=np.float32)
        for bc in ("dirichlet", "periodic"):
            st = heat_stencil(dims, 0.1, bc)
            for key in GPU_KEYS:
                if dims not in VARIANTS[key].dims or not availability(key)[0]:
                    continue
                var = VARIANTS[key]
                ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 9, var.storage,
                                            var.steps_per_pass)
                u = _run(key, u0, st, 9)
                err = np.abs(u - ref).max()
                assert err <= max(nr.TOLERANCE[var.storage], 2e-6), (key, bc, shape, err)


@pytest.mark.parametrize("key", ["hip-cdna4/zmarch", "hip-cdna4/ymarch", "hip-cdna4/naive"])
def test_odd_widths_and_vector_paths(key):
    """Widths that select 16-, 8- and 4-byte lane accesses and partial waves, both BCs."""
    dims = VARIANTS[key].dims[-1] if key.endswith("zmarch") else VARIANTS[key].dims[0]
    rng = np.random.default_rng(10)
    for nx in (3, 4, 6, 8, 64, 100, 255, 256, 257, 260, 516):
        shape = (9, 11, nx) if dims == 3 else (37, nx)
        u0 = rng.random(shape, dtype=np.float32)
        for bc in ("dirichlet", "periodic"):
            st = heat_stencil(dims, 0.9 / (2 * dims), bc)
            ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 7)
            _check(key, _run(key, u0, st, 7), ref)


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_naive_every_lane_width(storage):
    """1, 2 and 4 sites per thread (partial last tiles, both BCs) equal the oracle bit for bit."""
    _need("hip-cdna4/naive", 3)
    rng = np.random.default_rng(16)
    key = "hip-cdna4/naive" + ("" if storage == "float32" else "-bf16")
    for shape, bc in (((19, 37, 100), "periodic"), ((13, 70, 97), "dirichlet"),
                      ((5, 131, 517), "periodic"), ((37, 300), "dirichlet"), ((29, 257), "periodic")):
        u0 = rng.random(shape, dtype=np.float32)
        st = heat_stencil(len(shape), 0.9 / (2 * len(shape)), bc)
        ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 5, storage, 1)
        for vec in (1, 2, 4):
            s = make_stepper(key, u0, st)
            s.tile_override = _hip.Tile(0, 0, vec, 0, 0)
            print(f"[gpu] {key} {shape} {bc} vec={vec}", flush=True)
            try:
                s.step(5)
                u = s.get()
            finally:
                s.close()
            assert np.array_equal(u, ref), (shape, vec, nr.rel_err(u, ref))


def test_odd_widths_in_tuned_size_buckets():
    """Grids whose size bucket has tuned tiles but whose odd row width the tuned lane vector
    does not divide (121^3 beside the 128^3 entries) run on valid defaults, bit for bit."""
    rng = np.random.default_rng(18)
    for shape, bc in (((121, 121, 121), "dirichlet"), ((120, 124, 123), "periodic")):
        u0 = rng.random(shape, dtype=np.float32)
        st = heat_stencil(3, 0.15, bc)
        ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 12)
        for key in ("hip-cdna4/zmarch", "hip-cdna4/tb2", "hip-cdna4/tb3", "hip-cdna4/tb4"):
            _check(key, _run(key, u0, st, 12), ref)


@pytest.mark.parametrize("key,shape,bc", [
    ("hip-cdna4/zmarch", (40, 50, 64), "periodic"),
    ("hip-cdna4/tb3", (37, 41, 60), "dirichlet"),
    ("hip-cdna4/ymarch", (300, 260), "dirichlet"),
    ("hip-cdna4/tb8", (200, 130), "periodic"),
    ("hip-cdna4/tb4-bf16", (30, 40, 64), "periodic"),
])
def test_graph_replayed_passes_equal_direct_launches(key, shape, bc):
    """Small grids replay captured graphs of passes. Over several calls (host copies and
    synchronisations between the replays, buffer order flipped by odd pass counts) every
    result equals direct launches bit for bit, and the oracle for float32."""
    _need(key, len(shape))
    var = VARIANTS[key]
    k = var.steps_per_pass
    rng = np.random.default_rng(17)
    u0 = rng.random(shape, dtype=np.float32)
    st = heat_stencil(len(shape), 0.9 / (2 * len(shape)), bc)
    g = make_stepper(key, u0, st)
    d = make_stepper(key, u0, st)
    d.GRAPH_MAX_SITES = 0   # direct launches only
    # remainder steps shift the pass cadence, which only float32 results do not depend on
    extra = 1 if var.storage == "float32" and k > 1 else 0
    done = 0
    try:
        for n in (37 * k, 40 * k + extra, 17 * k, 18 * k):
            print(f"[gpu] graph {key} {shape} {bc} step({n})", flush=True)
            g.step(n)
            d.step(n)
            done += n
            ug, ud = g.get(), d.get()
            assert np.array_equal(ug, ud), (key, n, nr.rel_err(ug, ud))
            if var.storage == "float32":
                ref = nr.emulated_reference(u0, st.c0, st.c1, bc, done)
                assert np.array_equal(ug, ref), (key, n, nr.rel_err(ug, ref))
        assert len(g._graphs) >= 1 and not d._graphs
    finally:
        g.close()
        d.close()


ZMARCH_SHAPES = [(1, 16), (2, 8), (2, 4), (4, 4), (4, 8), (8, 2), (4, 2), (8, 1), (8, 8)]


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_zmarch_every_tile_shape(storage):
    """Every built (rows, waves) shape, vector width and store flag equals the oracle."""
    rng = np.random.default_rng(15)
    key = "hip-cdna4/zmarch" + ("" if storage == "float32" else "-bf16")
    for shape, bc in (((19, 37, 264), "periodic"), ((13, 70, 100), "dirichlet")):
        u0 = rng.random(shape, dtype=np.float32)
        st = heat_stencil(3, 0.15, bc)
        ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 5, storage, 1)
        vecs = (1, 2, 4) if storage == "float32" else (1, 2, 4, 8)
        for (r, w) in ZMARCH_SHAPES:
            for vec in vecs:
                if shape[2] % vec:
                    continue
                for zch, flags in ((0, 0), (7, 1)):
                    s = make_stepper(key, u0, st)
                    s.tile_override = _hip.Tile(r, zch, vec, w, flags)
                    print(f"[gpu] {key} {shape} {bc} rows={r} waves={w} vec={vec} zch={zch} "
                          f"flags={flags}", flush=True)
                    try:
                        s.step(5)
                        u = s.get()
                    finally:
                        s.close()
                    assert np.array_equal(u, ref), (r, w, vec, zch, nr.rel_err(u, ref))


TB3_SHAPES = [(2, 4, 16), (1, 8, 8), (2, 4, 8), (4, 2, 16), (1, 4, 16)]  # (vec, rows, waves)


@pytest.mark.parametrize("k", [2, 3, 4])
def test_tb3d_every_tile_shape(k):
    """Every built 3D temporal-blocking shape and lane vector (even widths), both BCs, source."""
    rng = np.random.default_rng(16 + k)
    for shape, bc in (((21, 67, 136), "periodic"), ((17, 40, 72), "dirichlet")):
        u0 = rng.random(shape, dtype=np.float32)
        b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
        st = heat_stencil(3, 0.15, bc)
        for src in (None, b):
            ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 2 * k, b=src)
            for (vec, r, w) in TB3_SHAPES:
                for zch in (0, 5):
                    s = make_stepper(f"hip-cdna4/tb{k}", u0, st, src)
                    s.tile_override = _hip.Tile(r, zch, vec, w, 0)
                    print(f"[gpu] tb{k} {shape} {bc} src={src is not None} vec={vec} rows={r} "
                          f"waves={w} zch={zch}", flush=True)
                    try:
                        s.step(2 * k)
                        u = s.get()
                    finally:
                        s.close()
                    assert np.array_equal(u, ref), (k, bc, vec, r, w, zch, nr.rel_err(u, ref))


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_tb4_16bit_every_tile_shape(storage):
    """16-bit tb4 in every built 3D shape (the periodic 2- and 4-site tiles run with pinned
    stages, the 16-wave 2-site and the sourced ones with one exchange per plane step), both
    BCs, with and without a float32 source, bit for bit against the storage-rounding oracle."""
    key = "hip-cdna4/tb4" + {"float16": "-fp16", "bfloat16": "-bf16"}[storage]
    rng = np.random.default_rng(44)
    for shape, bc in (((21, 67, 136), "periodic"), ((17, 40, 72), "dirichlet")):
        u0 = rng.random(shape, dtype=np.float32)
        b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
        st = heat_stencil(3, 0.15, bc)
        for src in (None, b):
            ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 8, storage, 4, src)
            for (vec, r, w) in TB3_SHAPES:
                s = make_stepper(key, u0, st, src)
                s.tile_override = _hip.Tile(r, 0, vec, w, 0)
                print(f"[gpu] {key} {shape} {bc} src={src is not None} vec={vec} rows={r} "
                      f"waves={w}", flush=True)
                try:
                    s.step(8)
                    u = s.get()
                finally:
                    s.close()
                assert np.array_equal(u, ref), (storage, bc, vec, r, w, nr.rel_err(u, ref))


@pytest.mark.parametrize("k", [2, 3, 4])
def test_tb3d_tiles_with_one_column_and_one_row_tails(k):
    """Grids one lane vector wider and one row taller than a tile's output, so the last tile of
    each row and column writes a single vector / row: lanes and rows past what that tile's
    outputs depend on load nothing, and the sites next to that cut must still come out bitwise,
    for every built shape, both BCs, with and without a source."""
    rng = np.random.default_rng(30 + k)
    for (vec, r, w) in TB3_SHAPES:
        kv = -(-k // vec) * vec
        ox, oy = 64 * vec - 2 * kv, w * r - 2 * k
        shape = (9, oy + 1, ox + vec)
        for bc in ("periodic", "dirichlet"):
            u0 = rng.random(shape, dtype=np.float32)
            b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
            st = heat_stencil(3, 0.15, bc)
            for src in (None, b):
                ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 2 * k, b=src)
                s = make_stepper(f"hip-cdna4/tb{k}", u0, st, src)
                s.tile_override = _hip.Tile(r, 0, vec, w, 0)
                print(f"[gpu] tb{k} tails {shape} {bc} src={src is not None} vec={vec} rows={r} "
                      f"waves={w}", flush=True)
                try:
                    s.step(2 * k)
                    u = s.get()
                finally:
                    s.close()
                assert np.array_equal(u, ref), (k, shape, bc, vec, r, w, nr.rel_err(u, ref))


@pytest.mark.parametrize("key", ["hip-cdna4/tb2", "hip-cdna4/tb3", "hip-cdna4/tb4",
                                 "hip-cdna4/tb8", "hip-cdna4/tb4-bf16"])
def test_tb2d_lane_vectors(key):
    """2D temporal blocking with every lane vector width (even and odd widths), both BCs, with
    and without a source."""
    rng = np.random.default_rng(30)
    k = VARIANTS[key].steps_per_pass
    storage = VARIANTS[key].storage
    for shape in ((37, 136), (50, 264), (19, 97)):
        u0 = rng.random(shape, dtype=np.float32)
        b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
        for bc in ("dirichlet", "periodic"):
            st = heat_stencil(2, 0.2, bc)
            for src in (None, b):
                ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 2 * k + 1, storage, k, src)
                for vec in (1, 2, 4, 8):
                    if shape[1] % vec or (vec == 8 and storage == "float32"):
                        continue
                    s = make_stepper(key, u0, st, src)
                    s.tile_override = _hip.Tile(0, 0, vec, 4, 0)
                    print(f"[gpu] {key} {shape} {bc} src={src is not None} vec={vec}", flush=True)
                    try:
                        s.step(2 * k + 1)
                        u = s.get()
                    finally:
                        s.close()
                    assert np.array_equal(u, ref), (key, shape, bc, vec, nr.rel_err(u, ref))


def test_tb8_odd_widths_take_site_by_site_two_site_lanes():
    """2D tb8 on widths that forbid vector lanes (odd, and 4-byte aligned only) runs 2-site
    lanes loading and storing site by site (the resolved launch says vec 2) and equals the
    oracle bit for bit, both BCs, with and without a source, on several strips and chunks."""
    _need("hip-cdna4/tb8", 2)
    rng = np.random.default_rng(31)
    for shape in ((131, 149), (70, 65), (300, 257), (41, 1001)):
        u0 = rng.random(shape, dtype=np.float32)
        b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
        for bc in ("dirichlet", "periodic"):
            st = heat_stencil(2, 0.2, bc)
            for src in (None, b):
                ref = nr.emulated_reference(u0, st.c0, st.c1, bc, 17, "float32", 8, src)
                s = make_stepper("hip-cdna4/tb8", u0, st, src)
                print(f"[gpu] tb8 {shape} {bc} src={src is not None} plan={s.tile}", flush=True)
                try:
                    assert s.tile["vec"] == 2, s.tile
                    s.step(17)
                    u = s.get()
                finally:
                    s.close()
                assert np.array_equal(u, ref), (shape, bc, src is not None, nr.rel_err(u, ref))


@pytest.mark.parametrize("bc", ["periodic", "dirichlet"])
def test_large_grid_every_3d_rung_matches_oracle(bc):
    """512^3, 12 steps (a multiple of every k): every 3D rung bit for bit against the scalar C
    oracle. Large enough that set-up copies take longer than a kernel launch, so a first sweep
    that is not ordered after the upload of both ping-pong buffers would show up here."""
    from oracle import cref

    n = 512
    u0 = np.random.default_rng(41).random((n, n, n), dtype=np.float32)
    st = heat_stencil(3, 0.15, bc)
    ref = cref.run_f32(u0, st.c0, st.c1, bc, 12)
    for key in ("hip-cdna4/naive", "hip-cdna4/zmarch", "hip-cdna4/tb2", "hip-cdna4/tb3",
                "hip-cdna4/tb4"):
        u = _run(key, u0, st, 12)
        bad = int(np.sum(u != ref))
        assert bad == 0, (key, bc, bad, nr.rel_err(u, ref))


@pytest.mark.parametrize("bc", ["periodic", "dirichlet"])
def test_full_size_1024cube_rungs_agree_and_match_oracle_windows(bc):
    """1024^3 float32 heat (r = 0.15), 10 steps, at the single-GPU benchmark size: every 3D rung
    returns the same 4.3 GB field bit for bit (tb2 runs 5 fused passes, tb3 3 and one single
    step, tb4 2 and two single steps), and windows of it at the corners, the centre and, for
    periodic fields, across the wrap-around in every axis, equal the oracle bit for bit. The
    full-field oracle runs too slowly on one core at this size; the windows use the light cone."""
    n, steps, size = 1024, 10, 32
    u0 = np.random.default_rng(1).random((n, n, n), dtype=np.float32)
    st = heat_stencil(3, 0.15, bc)
    lows = [(0, 0, 0), (n - size,) * 3, (n // 2 - size // 2,) * 3, (0, n // 2, n - size),
            (n - size, 0, n // 3)]
    if bc == "periodic":
        lows += [(n - size // 2,) * 3, (n - 3, 5, n - size // 4)]
    wins = [nr.window_reference(u0, lo, size, st.c0, st.c1, bc, steps) for lo in lows]
    first = None
    for key in ("hip-cdna4/naive", "hip-cdna4/zmarch", "hip-cdna4/tb2", "hip-cdna4/tb3",
                "hip-cdna4/tb4"):
        u = _run(key, u0, st, steps)
        for lo, (ref, idx) in zip(lows, wins):
            got = u[np.ix_(*idx)]
            assert np.array_equal(got, ref), (key, bc, lo, nr.rel_err(got, ref))
        if first is None:
            first = u
        else:
            bad = int(np.count_nonzero(u != first))
            assert bad == 0, (key, bc, bad, nr.rel_err(u, first))
        del u


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_full_size_768cube_jacobi_storage_matches_oracle_windows(storage):
    """768^3 Jacobi sweeps with a float32 source and 16-bit storage (the Poisson solve's size and
    settings), 10 sweeps from a random field: windows at the walls, the corners and the centre
    equal the storage oracle (rounding after every pass of k sweeps) bit for bit, for the naive,
    z-march and tb4 kernels; the two single-sweep kernels return the same field bit for bit."""
    n, steps, size = 768, 10, 32
    rng = np.random.default_rng(2)
    u0 = rng.random((n, n, n), dtype=np.float32)
    b = rng.random((n, n, n), dtype=np.float32) * np.float32(1e-3)
    st = jacobi_stencil(3)
    lows = [(0, 0, 0), (n - size,) * 3, (n // 2 - size // 2,) * 3, (0, n // 2, n - size)]
    suffix = {"float16": "-fp16", "bfloat16": "-bf16"}[storage]
    single = None
    for key in ("hip-cdna4/naive", "hip-cdna4/zmarch", "hip-cdna4/tb4"):
        key += suffix
        k = VARIANTS[key].steps_per_pass
        u = _run(key, u0, st, steps, b)
        for lo in lows:
            ref, idx = nr.window_reference(u0, lo, size, st.c0, st.c1, "dirichlet", steps,
                                           storage, k, b)
            got = u[np.ix_(*idx)]
            assert np.array_equal(got, ref), (key, lo, nr.rel_err(got, ref))
        if k == 1:
            if single is None:
                single = u
            else:
                bad = int(np.count_nonzero(u != single))
                assert bad == 0, (key, bad)
        del u


def test_residual_matches_oracle():
    rng = np.random.default_rng(21)
    for shape in ((33, 40, 47), (61, 75)):
        d = len(shape)
        u0 = rng.random(shape, dtype=np.float32)
        b = rng.random(shape, dtype=np.float32) * np.float32(1e-3)
        for bc in ("dirichlet", "periodic"):
            for key, src in (("hip-cdna4/naive", None), ("hip-cdna4/naive", b),
                             ("hip-cdna4/naive-bf16", b)):
                st = make_stepper(key, u0, jacobi_stencil(d, bc), src)
                try:
                    got = st.residual_sq()
                    stored = st.get()
                finally:
                    st.close()
                want = nr.residual_fp64(stored, 1 / (2 * d), bc, src)
                assert got == pytest.approx(want, rel=1e-12), (shape, bc, key)


def test_device_is_gfx950():
    info = _hip.device_info(0)
    print(info)
    assert info["arch"].startswith("gfx950")


def test_random_problems_every_variant():
    """Seeded random problems (shape 3..90 per axis with odd widths, boundary condition, step
    count, with or without a source, heat or Jacobi coefficients) for every variant: bitwise
    against the oracle with the variant's storage rounding."""
    rng = np.random.default_rng(2024)
    cases = 0
    for _ in range(12):
        dims = int(rng.integers(2, 4))
        shape = tuple(int(x) for x in rng.integers(3, 91, size=dims))
        bc = "periodic" if rng.random() < 0.5 else "dirichlet"
        steps = int(rng.integers(1, 18))
        u0 = rng.random(shape, dtype=np.float32)
        src = rng.random(shape, dtype=np.float32) * np.float32(1e-3) if rng.random() < 0.4 else None
        st = heat_stencil(dims, float(rng.uniform(0.02, 0.5 / dims)), bc) \
            if rng.random() < 0.7 else jacobi_stencil(dims, bc)
        for key in GPU_KEYS:
            var = VARIANTS[key]
            if dims not in var.dims:
                continue
            k = var.steps_per_pass if var.storage != "float32" else 1
            ref = nr.emulated_reference(u0, st.c0, st.c1, bc, steps, var.storage, k, src)
            u = _run(key, u0, st, steps, src)
            assert np.array_equal(u, ref), (key, shape, bc, steps, src is not None,
                                            nr.rel_err(u, ref))
            cases += 1
    assert cases > 100
