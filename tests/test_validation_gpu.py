"""``validate`` on the GPU: the reference's convergence studies run on float32 hip-cdna4 rungs.

The Laplace and Poisson studies use the reference's grids, tolerances and fixed sweep counts
(``validation.py:89-147``), so each run's sweep count equals the reference's, and its error
equals the reference's FP64 error (``results/validation.json`` via
``tests/golden/validation_golden.json``) up to what float32 does to the converged field. Jacobi
converges to the fixed point of its rounded map, which differs from the exact one by at most the
per-sweep rounding eps32 = 2^-24 amplified by the iteration's sensitivity 1 / (1 - cos(pi h)):
every run's error must lie that close to the reference's. The observed order must fall in the
reference's band (1.85, 2.15) over the grids where this float32 bound is below half the
discretisation error (2D: the bound reaches 2e-4 at n = 129, above the 1.7e-5 error, so float32
leaves the asymptotic range there; the C oracle gives the same float32 errors bit for bit).
"""

import json
import math
from pathlib import Path

import pytest

from pyheat3d_amd import validation as va

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "validation_golden.json").read_text())


def _golden(study: str, dims: int) -> dict:
    return next(x for x in GOLDEN[study] if x["dims"] == dims)


@pytest.mark.parametrize("study,dims,ns", [("laplace", 2, [9, 17, 33, 65, 129]),
                                           ("laplace", 3, [9, 17, 33, 65]),
                                           ("poisson", 2, [9, 17, 33, 65]),
                                           ("poisson", 3, [9, 17, 33])])
@pytest.mark.parametrize("key", ["hip-cdna4/naive", "hip-cdna4/march"])
def test_jacobi_order_studies_match_the_reference(study, dims, ns, key):
    if key == "hip-cdna4/march":
        key = "hip-cdna4/zmarch" if dims == 3 else "hip-cdna4/ymarch"
    fn = va.laplace_convergence if study == "laplace" else va.poisson_convergence
    print(f"[gpu] {study} {dims}D {key} {ns}", flush=True)
    res = fn(key, dims, ns)
    gold = _golden(study, dims)
    dominated = []
    for run, ref in zip(res["runs"], gold["runs"]):
        assert run["n"] == ref["n"] and run["iterations"] == ref["iterations"], (run, ref)
        bound = 2.0 ** -24 / (1.0 - math.cos(math.pi / (run["n"] - 1)))
        assert abs(run["err"] - ref["err"]) <= bound, (run, ref, bound)
        if bound < 0.5 * ref["err"]:
            dominated.append(run)
    assert len(dominated) >= 3, dominated
    order = va.slope([r["h"] for r in dominated], [r["err"] for r in dominated])
    print(f"[gpu] order {order:.4f} over n = {[r['n'] for r in dominated]} (reference "
          f"{gold['order']:.4f} over all); errors " + ", ".join(f"{r['err']:.3e}" for r in res["runs"]),
          flush=True)
    assert 1.85 < order < 2.15, order


def test_run_all_orders_in_band():
    res = va.run_all(quick=True)
    for x in res["spatial"] + res["laplace"] + res["poisson"]:
        assert 1.85 < x["order"] < 2.15, (x["variant"], x["dims"], x["order"])
    for x in res["temporal"]:
        assert 0.95 < x["order"] < 1.05, (x["variant"], x["order"])
    for x in res["laplace_quadratic"]:
        assert x["err"] < 1e-5, x
