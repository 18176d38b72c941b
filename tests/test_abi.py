"""The C ABI library: builds, loads without a GPU, exports every declared symbol. CPU only."""

import re
from pathlib import Path

import pytest

from pyheat3d_amd import _hip
from pyheat3d_amd.registry import VARIANTS, availability, hip_available

HEADER = Path(__file__).resolve().parents[1] / "include" / "pyheat3d_hip.h"


def declared() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return set(re.findall(r"\b(ph3_[a-z0-9_]+)\s*\(", text))


def test_header_declares_expected_entry_points():
    names = declared()
    for must in ("ph3_sweep", "ph3_residual", "ph3_encode", "ph3_decode", "ph3_copy_kernel",
                 "ph3_malloc", "ph3_event_elapsed_ms", "ph3_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _hip.lib()
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared() == set(_hip.SIGNATURES)


def test_error_path_without_gpu_is_an_exception_not_a_fallback():
    if _hip.device_count() > 0:
        pytest.skip("a GPU is present")
    assert not hip_available()
    ok, why = availability("hip-cdna4/naive")
    assert not ok and why
    # a launch with a bad shape fails with a message (argument checks run before any HIP call)
    with pytest.raises(_hip.HipError, match="at least 3 points"):
        _hip.sweep(_hip.NAIVE, 1, "float32", False, (2, 5), 0.2, 0.2, 16, 32, None, None, None)


def test_version_and_workspace():
    assert _hip.lib().ph3_version() >= 10000
    assert _hip.lib().ph3_residual_work_bytes() >= 8


def test_registry_keys_and_json():
    import json

    assert all(k.startswith("hip-cdna4/") for k in VARIANTS)
    json.dumps({k: v.description for k, v in VARIANTS.items()})
    assert VARIANTS["hip-cdna4/zmarch"].dims == (3,)
    assert VARIANTS["hip-cdna4/ymarch"].dims == (2,)


def test_header_constants_match_the_binding():
    """The Python mirrors of the header's #defines (kind codes, batch constants) agree."""
    defines = {m.group(1): int(m.group(2)) for m in
               re.finditer(r"#define (PH3_[A-Z_]+) \(?(-?\d+)\)?", HEADER.read_text())}
    assert (defines["PH3_NAIVE"], defines["PH3_ZMARCH"], defines["PH3_TB"]) == (
        _hip.NAIVE, _hip.ZMARCH, _hip.TB)
    assert defines["PH3_BATCH_CHECK"] == _hip.BATCH_CHECK
    assert defines["PH3_BATCH_SMALL_SITES"] == _hip.BATCH_SMALL_SITES
    assert defines["PH3_STALL_RING"] == _hip.STALL_RING
