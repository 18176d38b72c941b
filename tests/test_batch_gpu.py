"""Batched solves on the GPU: parity with the oracle, alone == in the mix, batch == one by one.

The mix spans both tiers (workgroup-owned grids of <= 4096 sites and varlen tile passes),
2D and 3D, heat and Jacobi, float32 and bfloat16.
"""

import numpy as np
import pytest

from oracle import numpy_ref as nr
from pyheat3d_amd import _hip
from pyheat3d_amd.analytic import splitmix_field
from pyheat3d_amd.batch import BatchSolver, Request, solve_one_by_one

pytestmark = pytest.mark.gpu


def mix(storage="float32"):
    shapes = [(20, 20), (33, 47), (130, 150), (200, 201), (12, 13, 14), (40, 36, 30),
              (9, 70, 70), (5, 6, 7)]
    reqs = []
    for i, sh in enumerate(shapes):
        for mode in ("heat", "jacobi"):
            # 40 heat steps let float32 large grids fuse 8 (2D) / 2 (3D) steps per pass; 37 do not
            heat_steps = 40 if i % 2 else 37
            reqs.append(Request(shape=sh, mode=mode, storage=storage,
                                steps=heat_steps if mode == "heat" else 384, seed=1000 + 17 * i,
                                rid=len(reqs)))
    # float32 3D grids of >= 128^3 sites fuse 4 steps per pass; (9, 70, 70) Jacobi above takes
    # the 2-site-lane class of 3D 2-step grids of even width
    reqs.append(Request(shape=(130, 128, 128), mode="heat", storage=storage, steps=40, seed=999,
                        rid=len(reqs)))
    return reqs


def run(reqs, graph=True, round_sweeps=128):
    print(f"[gpu] batch of {len(reqs)} graph={graph}", flush=True)
    b = BatchSolver(reqs, round_sweeps=round_sweeps, use_graph=graph)
    try:
        b.run()
        return b.results(keep_fields=True)
    finally:
        b.close()


def test_fill_random_matches_host_mirror():
    stream = _hip.Stream()
    for shape, kind in (((7, 9, 11), 0), ((13, 5), 1), ((3, 4, 5), 2)):
        nz, ny, nx = (1, *shape) if len(shape) == 2 else shape
        p = _hip.malloc(int(np.prod(shape)) * 4)
        try:
            _hip.check(_hip.lib().ph3_fill_random(0, p, len(shape), nz, ny, nx, 99, kind,
                                                  stream.handle))
            stream.sync()
            got = np.empty(shape, np.float32)
            _hip.memcpy(got.ctypes.data, p, got.nbytes, _hip.D2H)
        finally:
            _hip.free(p)
        assert np.array_equal(got, splitmix_field(shape, 99, kind))
    stream.close()


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_batch_matches_oracle(storage):
    reqs = mix(storage)
    res = run(reqs)
    for req, r in zip(reqs, res):
        u0 = splitmix_field(req.shape, req.seed, req.kind)
        st = req.stencil()
        assert r.status in ("done", "stagnated", "max_sweeps"), (req, r.status)
        if req.mode == "heat":
            assert r.status == "done" and r.steps == req.steps
        ref = nr.emulated_reference(u0, st.c0, st.c1, "dirichlet", r.steps, storage, 1)
        assert np.array_equal(r.u, ref), (req.shape, req.mode, r.steps, nr.rel_err(r.u, ref))
        if req.mode == "jacobi":
            # the relative residual reported by the device equals the oracle's FP64 residual
            r0 = nr.residual_fp64(nr.round_storage(u0, storage), st.c1, "dirichlet")
            rel = np.sqrt(nr.residual_fp64(ref, st.c1, "dirichlet") / r0)
            assert r.rel_residual == pytest.approx(rel, rel=1e-9)
            if r.status == "done":
                assert r.rel_residual <= req.tol


def test_periodic_heat_matches_oracle():
    """Periodic solves in the varlen passes (fused and single steps, 2D rows of odd and even
    width, 3D with 1- and 2-site lanes), next to Dirichlet ones of the same shapes."""
    shapes = [(200, 201), (190, 256), (300, 129), (40, 36, 30), (33, 34, 35), (20, 24, 100)]
    reqs = []
    for i, sh in enumerate(shapes):
        for bc in ("periodic", "dirichlet"):
            for steps in (40, 37):
                reqs.append(Request(shape=sh, mode="heat", steps=steps, seed=77 + 5 * i, bc=bc,
                                    rid=len(reqs)))
    res = run(reqs)
    for req, r in zip(reqs, res):
        u0 = splitmix_field(req.shape, req.seed, req.kind)
        st = req.stencil()
        assert r.status == "done" and r.steps == req.steps
        ref = nr.emulated_reference(u0, st.c0, st.c1, req.bc, req.steps, "float32", 1)
        assert np.array_equal(r.u, ref), (req.shape, req.bc, req.steps, nr.rel_err(r.u, ref))


def test_alone_equals_in_mix():
    reqs = mix()
    together = run(reqs)
    for req, r in zip(reqs, together):
        alone = run([req])[0]
        assert alone.status == r.status and alone.steps == r.steps, req
        assert np.array_equal(alone.u, r.u), req


def test_batch_equals_one_by_one_and_graph_equals_eager():
    reqs = mix()
    a = run(reqs, graph=True)
    b = run(reqs, graph=False, round_sweeps=64)
    c, _ = solve_one_by_one(reqs, keep_fields=True)
    for req, x, y, z in zip(reqs, a, b, c):
        assert x.steps == y.steps == z.steps, (req, x.steps, y.steps, z.steps)
        assert np.array_equal(x.u, y.u) and np.array_equal(x.u, z.u), req


def test_jacobi_converges_on_small_grid():
    req = Request(shape=(17, 17), mode="jacobi", steps=100000, seed=5)
    r = run([req])[0]
    assert r.status == "done" and r.rel_residual <= 1e-6
    predicted = np.log(1e-6) / np.log(np.cos(np.pi / 16))
    assert r.steps < 3 * predicted


def test_per_solve_on_the_batch_rung_equals_the_batch():
    """The comparator that runs each solve on the fused-step rung the batch chose (tb<kpass>)
    gives the batch's bits, step counts and statuses, and that rung is really fused (kpass > 1
    for the float32 tier-2 solves of the mix)."""
    from pyheat3d_amd.batch import batch_kpass

    reqs = mix()
    b = BatchSolver(reqs, round_sweeps=128)
    try:
        b.run()
        kp = batch_kpass(b)
        batched = b.results(keep_fields=True)
    finally:
        b.close()
    assert any(k > 1 for k in kp.values()), kp
    per, _ = solve_one_by_one(reqs, keep_fields=True, kpass=kp)
    for req, x, y in zip(reqs, batched, per):
        assert (x.status, x.steps) == (y.status, y.steps), (req.shape, req.mode, kp[req.rid])
        assert np.array_equal(x.u, y.u), (req.shape, req.mode, kp[req.rid])


def test_results_against_fp64_solutions():
    """Heat results within the reference's float32 tolerance of the FP64 replay; a Jacobi solve
    that converged to 1e-6 within 1e-5 of the exact FP64 fixed point; completion times set."""
    from pyheat3d_amd.batch import error_vs_fp64

    reqs = [Request(shape=(130, 150), mode="heat", steps=200, seed=41, rid=0),
            Request(shape=(40, 36, 30), mode="heat", steps=200, seed=42, rid=1),
            Request(shape=(17, 17), mode="jacobi", steps=100000, seed=43, rid=2),
            Request(shape=(12, 13, 14), mode="jacobi", steps=100000, seed=44, rid=3)]
    b = BatchSolver(reqs, round_sweeps=256)
    try:
        b.run()
        res = b.results(keep_fields={0, 1, 2, 3})
    finally:
        b.close()
    for q, r in zip(reqs, res):
        assert r.seconds is not None and r.seconds > 0
        e = error_vs_fp64(q, r.u)
        print(f"[gpu] {q.shape} {q.mode} {r.status} error vs FP64 {e:.3e}", flush=True)
        if q.mode == "heat":
            assert e <= nr.TOLERANCE["float32"], (q.shape, e)
        else:
            assert r.status == "done" and e <= 1e-5, (q.shape, r.status, e)


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6])
def test_every_tier2_tile_shape_and_chunk_policy_matches_oracle(tile):
    """Each fused tile shape (ph3_solve_desc.tile; 2-site shapes fall back to 1-site lanes on
    odd widths) under non-default chunk policies: 2D 8-step, 3D 2-step and 4-step passes, both
    boundaries, and a Poisson solve with a source, bit for bit against the oracle."""
    from oracle import cref

    rng = np.random.default_rng(60 + tile)
    reqs = []
    for sh in ((130, 150), (131, 149), (40, 36, 70), (33, 34, 67), (130, 128, 128)):
        for bc in ("dirichlet", "periodic"):
            # 3D-only shape 6: 2D solves take their library default
            t2 = tile if len(sh) == 3 or tile <= 5 else 0
            reqs.append(Request(shape=sh, mode="heat", steps=40, bc=bc, seed=700 + len(reqs),
                                rid=len(reqs), tile=t2))
    src = (rng.random((40, 36, 70), dtype=np.float32) * np.float32(1e-3)).astype(np.float32)
    reqs.append(Request(shape=(40, 36, 70), mode="jacobi", steps=256, seed=777, source=src,
                        rid=len(reqs), tile=tile))
    params = [(2, 8, 4096, 256), (3, 2, 8192, 128), (3, 4, 2048, 128)]
    print(f"[gpu] tile shape {tile}: {len(reqs)} solves", flush=True)
    b = BatchSolver(reqs, round_sweeps=128, params=params, tuned_tiles=False)
    try:
        b.run()
        res = b.results(keep_fields=True)
        kp = {q.rid: int(d.kpass) for q, d in zip(reqs, b.state)}
    finally:
        b.close()
    assert {kp[q.rid] for q in reqs if q.dims == 2} == {8}
    assert {kp[q.rid] for q in reqs if q.dims == 3} == {2, 4}
    for q, r in zip(reqs, res):
        st = q.stencil()
        ref = cref.run_f32(splitmix_field(q.shape, q.seed, q.kind), st.c0, st.c1, q.bc, r.steps,
                           q.source)
        assert np.array_equal(r.u, ref), (tile, q.shape, q.bc, q.mode, nr.rel_err(r.u, ref))


def test_set_rejects_planes_beyond_the_buffer_window():
    """A tier-2 solve whose plane of float32 values exceeds the buffer resources' 2 GiB window
    is refused by ph3_batch_set before anything touches its buffers (the pointers here are never
    dereferenced), and the batch keeps working for a valid table afterwards."""
    from pyheat3d_amd.batch import Batch

    stream = _hip.Stream()
    b = Batch(2, stream)
    try:
        d = _hip.SolveDesc()
        d.dims, d.nz, d.ny, d.nx = 3, 3, 32768, 32768      # a 4 GiB float32 plane
        d.storage, d.c0, d.c1, d.mode, d.steps = 0, 0.0, 1.0 / 6.0, 0, 4
        d.buf[0], d.buf[1] = 0x100000, 0x200000
        with pytest.raises(_hip.HipError, match="buffer window"):
            b.set([d])
        req = Request(shape=(40, 36, 70), mode="heat", steps=8, seed=3)
        s = BatchSolver([req], tuned_tiles=False)
        try:
            s.run()
            assert s.results()[0].status == "done"
        finally:
            s.close()
    finally:
        b.close()
        stream.close()
