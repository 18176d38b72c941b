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


def test_oversized_request_is_rejected():
    srv = Server(10 ** 5)
    try:
        with pytest.raises(ValueError, match="HBM budget"):
            srv.serve([Request(shape=(200, 200))], np.array([0.0]))
    finally:
        srv.close()
