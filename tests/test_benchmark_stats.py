"""Outcome statistics and accuracy sampling of the batch / serve benchmarks (CPU only)."""

import pytest

from pyheat3d_amd.batch import Request, Result
from pyheat3d_amd.benchmark import accuracy_sample, outcome_stats


def test_outcome_stats_separate_reached_from_stagnated():
    reqs = [Request(shape=(70, 70), mode="heat", rid=0),
            Request(shape=(70, 70), mode="jacobi", rid=1),
            Request(shape=(70, 70), mode="jacobi", rid=2),
            Request(shape=(70, 70), mode="jacobi", rid=3)]
    res = [Result(0, "done", 200, None, seconds=0.1),
           Result(1, "done", 4096, 9e-7, seconds=0.3),
           Result(2, "stagnated", 8192, 2e-5, seconds=0.9),
           Result(3, "max_sweeps", 16384, 3e-4, seconds=1.0)]
    st = outcome_stats(reqs, res, 2.0)
    assert st["solves"] == 4 and st["solves_per_s"] == pytest.approx(2.0)
    assert st["reached_target"]["solves"] == 2
    assert st["reached_target"]["solves_per_s"] == pytest.approx(1.0)
    assert st["reached_target"]["max_ms"] == pytest.approx(300.0)
    assert st["stagnated"]["solves"] == 1 and st["max_sweeps"]["solves"] == 1
    assert st["jacobi"] == {"solves": 3, "converged": 1, "stagnated": 1, "capped": 1,
                            "max_rel_residual": 3e-4}


def test_accuracy_sample_is_systematic_and_bounded():
    reqs = [Request(shape=(2048, 2048), mode="jacobi", rid=0),
            Request(shape=(256, 256, 256), mode="jacobi", rid=1),
            Request(shape=(256, 256, 256), mode="heat", steps=200, rid=2),
            Request(shape=(64, 64), mode="heat", rid=3)]
    assert accuracy_sample(reqs, every=1) == {0, 3}   # 256^3: too large for the FP64 solutions
    assert accuracy_sample(reqs, every=2) == {0}
