"""The benchmark results file keeps the reference's record schema.

``results/mi355x_dev.json`` is the output of ``python -m pyheat3d_amd bench --dev`` on an MI355X.
The reference's figure and paper-macro script reads benchmark files through the fields checked
here (``figures/make_figures.py:44-95`` loading and record helpers, ``:298-372`` macros and the
performance table; records written by ``src/pyheat3d/benchmark.py:189-233``), so a file that
passes can be dropped into its ``results/`` folder. The one addition a maintainer makes there
is a colour for the ``hip-cdna4`` backend, since the table orders rows by its colour list.
"""

import json
import math
from pathlib import Path

import pytest

RESULTS = Path(__file__).resolve().parents[1] / "results" / "mi355x_dev.json"

# fields the figure script indexes on every benchmark record
RECORD_FIELDS = ("backend", "variant", "key", "rung", "dims", "size", "dtype", "median", "min",
                 "max", "steps_per_pass", "bytes_per_unit", "effective_GBs", "pct_copy_bw",
                 "pct_spec_bw", "flops_per_unit", "GFLOPs")


@pytest.fixture(scope="module")
def bench():
    return json.loads(RESULTS.read_text())


def test_top_level_fields(bench):
    assert isinstance(bench["machine"], dict)
    for k in ("gpu", "date", "driver"):
        assert bench["machine"].get(k), k
    ceil = bench["ceiling"]
    for k in ("copy_bandwidth_GBs", "spec_bandwidth_GBs", "measured_copy_bandwidth_GBs"):
        assert isinstance(ceil[k], float) and ceil[k] > 0, k
    assert len(bench["runs"]) > 0


def test_every_record_has_the_read_fields(bench):
    for r in bench["runs"]:
        missing = [k for k in RECORD_FIELDS if k not in r]
        assert not missing, (r.get("key"), missing)
        assert r["backend"] == "hip-cdna4"
        assert len(r["size"]) == r["dims"]
        if r["steps_per_pass"] > 1:   # temporal-blocking rows also print this column
            assert r["single_step_equiv_GBs"] > 0


def test_record_arithmetic_is_consistent(bench):
    """Derived columns follow from the median and the traffic model (benchmark.py:184-212)."""
    copy_bw = bench["ceiling"]["measured_copy_bandwidth_GBs"]
    for r in bench["runs"]:
        k, bpv = r["steps_per_pass"], {"float32": 4, "float16": 2, "bfloat16": 2}[r["dtype"]]
        assert r["bytes_per_unit"] == pytest.approx(2.0 * bpv / k)
        assert r["effective_GBs"] == pytest.approx(r["median"] * r["bytes_per_unit"] / 1e3)
        assert r["pct_copy_bw"] == pytest.approx(100.0 * r["effective_GBs"] / copy_bw)
        assert r["pct_spec_bw"] == pytest.approx(100.0 * r["effective_GBs"] / 8000.0)
        assert r["flops_per_unit"] == 2 * r["dims"] + 2
        assert r["min"] <= r["median"] <= r["max"]
        assert r["finite"] is True
        assert math.isfinite(r["median"]) and r["median"] > 0


def test_covers_every_rung_at_every_size(bench):
    """Each GPU variant appears at each size of the dev plan that fits its dimensionality."""
    from pyheat3d_amd.benchmark import PLANS
    from pyheat3d_amd.registry import VARIANTS

    seen = {(r["key"], tuple(r["size"])) for r in bench["runs"]}
    for key, var in VARIANTS.items():
        if not var.gpu:
            continue
        for dims in var.dims:
            for n in PLANS["dev"]["sizes"][dims]:
                assert (key, (n,) * dims) in seen, (key, n, dims)


def test_records_carry_tile_gpu_state_and_halo_model(bench):
    """The reference's per-record tile, gpu_state and (temporal blocking) halo-model fields
    (benchmark.py:216-231): the tile is the launch the library resolved (ph3_sweep_plan)."""
    assert bench["machine"].get("cpu")
    for r in bench["runs"]:
        assert isinstance(r["gpu_state"], dict), r["key"]
        t = r["tile"]
        for f in ("vec", "waves", "zch", "tile_x", "out_x", "threads", "groups"):
            assert f in t, (r["key"], f)
        assert t["groups"] > 0 and t["out_x"] <= t["tile_x"]
        if r["steps_per_pass"] > 1:
            assert r["halo_model_bytes_per_unit"] > r["bytes_per_unit"], r["key"]
            assert r["halo_model_GBs"] == pytest.approx(
                r["median"] * r["halo_model_bytes_per_unit"] / 1e3)


def test_halo_model_of_a_plan_is_the_reference_formula():
    """With a halo of exactly K sites per side, the plan-based model is traffic.tb_model_bytes."""
    from pyheat3d_amd import traffic

    for dims, K, tx, ty, zch in ((3, 4, 64, 64, 128), (3, 2, 64, 32, 64), (2, 8, 64, 1, 128)):
        plan = {"tile_x": tx, "out_x": tx - 2 * K, "tile_y": ty,
                "out_y": ty - 2 * K if dims == 3 else 1, "zch": zch}
        assert traffic.halo_model_bytes(dims, K, 4, plan) == pytest.approx(
            traffic.tb_model_bytes(dims, K, 4, tx, ty, zch))
