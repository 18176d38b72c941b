"""Batched, served and single-grid solves at the workloads' full sizes against the oracle.

* Heat requests at 2D 2048^2 and 3D 256^3 (200 steps, Dirichlet and periodic), the largest
  grids of the mixed batch workload, plus Jacobi Laplace solves to 1e-6 at the largest sizes
  the scalar C oracle replays in seconds: batched and served, bit for bit against the oracle,
  and the device's FP64 residual against the oracle's.
* Poisson (source) requests in float32, float16 and bfloat16 in both batch tiers, including
  the fused-step and two-site-lane classes.
* The round graph across table compactions (the table is rewritten while a graph recorded
  against it exists).
* The 768^3 Poisson solve in float32 with the 4-step temporal-blocking kernel and a float32
  source, from zero on the eigen-source: the FP64 relative residual equals mu^s, with
  mu = 6 c1 cos(pi h) for the float32 coefficient c1 (the Jacobi operator maps the mode v to
  mu v, so the residual of u_s is mu^s b), and light-cone windows equal the oracle bit for bit.
"""

import math

import numpy as np
import pytest

from oracle import cref
from oracle import numpy_ref as nr
from pyheat3d_amd.analytic import poisson_eigen_source, splitmix_field
from pyheat3d_amd.batch import BatchSolver, Request
from pyheat3d_amd.problem import jacobi_stencil
from pyheat3d_amd.registry import make_stepper
from pyheat3d_amd.server import Server, poisson_arrivals

pytestmark = pytest.mark.gpu

_ORACLE: dict = {}


def oracle_field(req: Request, steps: int) -> np.ndarray:
    """The oracle's state of a request after ``steps`` steps (cached per request and count)."""
    key = (req.shape, req.mode, req.storage, req.seed, req.bc, req.source is not None, steps)
    if key not in _ORACLE:
        u0 = splitmix_field(req.shape, req.seed, req.kind)
        st = req.stencil()
        if req.storage == "float32":
            _ORACLE[key] = cref.run_f32(u0, st.c0, st.c1, req.bc, steps, req.source)
        else:
            _ORACLE[key] = nr.emulated_reference(u0, st.c0, st.c1, req.bc, steps, req.storage,
                                                 1, req.source)
    return _ORACLE[key]


def check_against_oracle(reqs, results) -> None:
    for req, r in zip(reqs, results):
        assert r.status in ("done", "stagnated", "max_sweeps"), (req.shape, r.status)
        if req.mode == "heat":
            assert r.status == "done" and r.steps == req.steps, (req.shape, r.status, r.steps)
        ref = oracle_field(req, r.steps)
        assert np.array_equal(r.u, ref), (req.shape, req.mode, req.bc, req.storage, r.steps,
                                          nr.rel_err(r.u, ref))
        if req.mode == "jacobi":
            st = req.stencil()
            u0 = nr.round_storage(splitmix_field(req.shape, req.seed, req.kind), req.storage)
            r0 = nr.residual_fp64(u0, st.c1, "dirichlet", req.source)
            rel = math.sqrt(nr.residual_fp64(ref, st.c1, "dirichlet", req.source) / r0)
            assert r.rel_residual == pytest.approx(rel, rel=1e-9), (req.shape, req.storage)
            if r.status == "done":
                assert r.rel_residual <= req.tol


def full_size_requests() -> list[Request]:
    reqs = []
    for sh in ((2048, 2048), (256, 256, 256)):
        for bc in ("dirichlet", "periodic"):
            reqs.append(Request(shape=sh, mode="heat", steps=200, bc=bc, seed=4000 + len(reqs),
                                rid=len(reqs)))
    # Jacobi Laplace to 1e-6 (random boundary values, zero inside): ~3e4 and ~6e3 sweeps
    for sh in ((128, 130), (48, 50, 52)):
        reqs.append(Request(shape=sh, mode="jacobi", steps=200000, seed=4100 + len(reqs),
                            rid=len(reqs)))
    return reqs


def test_batch_full_size_matches_oracle():
    reqs = full_size_requests()
    print(f"[gpu] full-size batch of {len(reqs)}", flush=True)
    b = BatchSolver(reqs, round_sweeps=256)
    try:
        b.run()
        res = b.results(keep_fields=True)
    finally:
        b.close()
    check_against_oracle(reqs, res)
    for req, r in zip(reqs, res):
        if req.mode == "jacobi":
            assert r.steps > 1000, (req.shape, r.steps)   # a real solve, not a short stop


def test_server_full_size_matches_oracle():
    reqs = full_size_requests()
    print(f"[gpu] full-size server of {len(reqs)}", flush=True)
    srv = Server(10 ** 10, round_sweeps=64, keep_fields=True)
    try:
        res, stats = srv.serve(reqs, poisson_arrivals(len(reqs), 1000.0, seed=9),
                               time_limit=300)
    finally:
        srv.close()
    assert stats.solves == len(reqs)
    check_against_oracle(reqs, res)


def poisson_requests() -> list[Request]:
    """Jacobi with a float32 source in every storage type and both tiers: tier-1 grids (at
    most 4096 sites), tier-2 2D (float32: 8-step passes), tier-2 3D with 2-site lanes (float32,
    even width above 64) and 1-site lanes."""
    rng = np.random.default_rng(31)
    reqs = []
    for storage in ("float32", "float16", "bfloat16"):
        for sh in ((40, 50), (12, 14, 16), (130, 140), (20, 30, 70), (33, 34, 35)):
            src = (rng.random(sh, dtype=np.float32) * np.float32(1e-3)).astype(np.float32)
            reqs.append(Request(shape=sh, mode="jacobi", storage=storage, steps=2048,
                                seed=5000 + len(reqs), source=src, rid=len(reqs)))
        # heat beside them in the same storage (200 steps: float32 fuses 8 / 2 steps)
        reqs.append(Request(shape=(150, 160), mode="heat", storage=storage, steps=200,
                            seed=5000 + len(reqs), rid=len(reqs)))
    return reqs


def test_poisson_requests_every_storage_both_tiers():
    reqs = poisson_requests()
    print(f"[gpu] poisson batch of {len(reqs)}", flush=True)
    b = BatchSolver(reqs, round_sweeps=128)
    try:
        b.run()
        res = b.results(keep_fields=True)
    finally:
        b.close()
    check_against_oracle(reqs, res)
    srv = Server(10 ** 9, round_sweeps=64, keep_fields=True)
    try:
        served, _ = srv.serve(reqs, poisson_arrivals(len(reqs), 2000.0, seed=10), time_limit=120)
    finally:
        srv.close()
    for req, x, y in zip(reqs, res, served):
        assert (x.status, x.steps) == (y.status, y.steps), req.shape
        assert np.array_equal(x.u, y.u), (req.shape, req.storage)


def test_round_graph_across_compactions():
    """Solves that finish at different rounds: the table is compacted (rewritten) several times
    while graphs recorded against earlier tables exist; every result still equals the oracle."""
    reqs = []
    for i, steps in enumerate((64, 128, 256, 384, 512, 640, 768, 1024)):
        for sh in ((70, 90), (30, 31, 32), (20, 20)):
            reqs.append(Request(shape=sh, mode="heat", steps=steps, seed=6000 + len(reqs),
                                rid=len(reqs)))
    b = BatchSolver(reqs, round_sweeps=64, use_graph=True, compact=True)
    try:
        timing = b.run()
        res = b.results(keep_fields=True)
    finally:
        b.close()
    assert timing["compactions"] >= 2, timing
    check_against_oracle(reqs, res)


def test_full_size_768cube_poisson_fp32_tb4_closed_form():
    n = 768
    v, b, _ = poisson_eigen_source(n, 3)
    del v
    b32 = b.astype(np.float32)
    del b
    u0 = np.zeros(b32.shape, np.float32)
    c1 = float(np.float32(1 / 6))
    mu = 6 * c1 * math.cos(math.pi / (n - 1))
    print("[gpu] 768^3 Poisson fp32 tb4 with source", flush=True)
    st = make_stepper("hip-cdna4/tb4", u0, jacobi_stencil(3), b32)
    try:
        r0 = st.residual_sq()
        assert r0 == pytest.approx(nr.residual_fp64(u0, c1, "dirichlet", b32), rel=1e-12)
        for s in (12, 40, 80, 120, 160, 200):
            st.step(s - st.steps_done)
            rel = math.sqrt(st.residual_sq() / r0)
            assert rel == pytest.approx(mu ** s, rel=1e-6), (s, rel, mu ** s)
            if s in (12, 200):
                u = st.get()
                if s == 12:   # centre, a face and an edge: light cones of 12 sites
                    lows = [(n // 2 - 16,) * 3, (0, n // 2, n // 2), (n - 32, 0, n // 3)]
                else:         # corners: light cones clipped by three walls
                    lows = [(0, 0, 0), (n - 32,) * 3]
                for lo in lows:
                    ref, idx = nr.window_reference(u0, lo, 32, 0.0, 1 / 6, "dirichlet", s,
                                                   b=b32, engine="c")
                    got = u[np.ix_(*idx)]
                    assert np.array_equal(got, ref), (s, lo, nr.rel_err(got, ref))
                del u
    finally:
        st.close()
