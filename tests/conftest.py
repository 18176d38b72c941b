"""Test configuration: the ``gpu`` marker and the repo root on ``sys.path``.

GPU tests are skipped only when the machine has no AMD GPU at all (no ``/dev/kfd``). On a
machine with a GPU they run and fail loudly if ``libpyheat3d_hip.so`` is missing or broken:
there is no CPU fallback to pass on.
"""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def has_gpu() -> bool:
    return os.path.exists("/dev/kfd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


def pytest_collection_modifyitems(config, items):
    if has_gpu():
        return
    skip = pytest.mark.skip(reason="no AMD GPU (/dev/kfd absent)")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
