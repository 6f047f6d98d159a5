"""Shared test setup: import paths and the `gpu` marker."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "graph-tool-rs_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgt_hip.so)")


def _have_gpu() -> bool:
    try:
        import gt_hip
        return gt_hip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a GPU test run without a device is an error, not a skip."""
    if not _have_gpu():
        pytest.fail("no HIP device visible (or libgt_hip.so missing)")
    return True
