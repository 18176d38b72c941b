"""Continuous-batching server on the GPU: admission by budget, results identical alone and mixed."""

import numpy as np
import pytest

from pyheat3d_amd.batch import BatchSolver, Request
from pyheat3d_amd.server import Server, poisson_arrivals

pytestmark = pytest.mark.gpu


def requests():
    shapes = [(24, 24), (130, 260), (10, 12, 14), (36, 36, 36), (64, 64), (300, 150),
              (7, 9, 11), (50, 40, 30), (96, 33)]
    out = []
    for i, sh in enumerate(shapes * 2):
        mode = "heat" if i % 2 else "jacobi"
        out.append(Request(shape=sh, mode=mode, steps=60 if mode == "heat" else 2000,
                           seed=77 + i, rid=i, storage="bfloat16" if i % 3 == 0 else "float32"))
    return out


def alone(req):
    b = BatchSolver([req], round_sweeps=128)
    try:
        b.run()
        return b.results(keep_fields=True)[0]
    finally:
        b.close()


@pytest.mark.parametrize("budget", [10 ** 9, 2 * 10 ** 6])
def test_server_results_equal_alone(budget):
    reqs = requests()
    arr = poisson_arrivals(len(reqs), rate=2000.0, seed=3)
    print(f"[gpu] server budget={budget} requests={len(reqs)}", flush=True)
    srv = Server(budget, round_sweeps=64, keep_fields=True)
    try:
        res, stats = srv.serve(reqs, arr, time_limit=120)
    finally:
        srv.close()
    assert stats.solves == len(reqs)
    assert stats.max_resident_bytes <= budget
    if budget < 10 ** 7:
        assert stats.max_resident < len(reqs)  # the budget forced queuing
    assert 0 < stats.p50_ms <= stats.p99_ms <= stats.max_ms
    for req, r in zip(reqs, res):
        a = alone(req)
        assert (a.status, a.steps) == (r.status, r.steps), req
        assert np.array_equal(a.u, r.u), req
        assert r.seconds >= 0


def test_served_heat_accuracy_vs_fp64():
    """Served heat solves against the FP64 oracle (same initial field, float64 arithmetic):
    float32 within the reference's cross-backend tolerance, bfloat16 within its storage error."""
    from oracle import numpy_ref as nr
    from pyheat3d_amd.analytic import splitmix_field

    shapes = [(130, 260), (36, 36, 36), (300, 150), (50, 40, 30)]
    reqs = [Request(shape=sh, mode="heat", steps=100, seed=300 + i, rid=2 * i + j,
                    storage=st)
            for i, sh in enumerate(shapes) for j, st in enumerate(("float32", "bfloat16"))]
    srv = Server(10 ** 9, round_sweeps=64, keep_fields=True)
    try:
        res, _ = srv.serve(reqs, poisson_arrivals(len(reqs), rate=500.0, seed=4), time_limit=120)
    finally:
        srv.close()
    worst = {"float32": 0.0, "bfloat16": 0.0}
    for req, r in zip(reqs, res):
        st = req.stencil()
        c0, c1 = float(np.float32(st.c0)), float(np.float32(st.c1))  # the device's coefficients
        ref = splitmix_field(req.shape, req.seed, req.kind).astype(np.float64)
        for _ in range(req.steps):
            ref = nr.reference_step(ref, c0, c1, "dirichlet")
        worst[req.storage] = max(worst[req.storage], nr.rel_err(r.u, ref))
    print(f"[gpu] served heat vs FP64 oracle, max relative error: {worst}", flush=True)
    assert worst["float32"] <= nr.TOLERANCE["float32"]
    assert worst["bfloat16"] <= 0.1


def test_oversized_request_is_rejected():
    srv = Server(10 ** 5)
    try:
        with pytest.raises(ValueError, match="HBM budget"):
            srv.serve([Request(shape=(200, 200))], np.array([0.0]))
    finally:
        srv.close()
