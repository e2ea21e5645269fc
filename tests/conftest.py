import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_report_header(config):
    return f"brr tests: root={ROOT}"


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with a visible device; fails loudly if either is missing."""
    import bayesrrcmd_amd as brr
    L = brr.lib()
    import ctypes
    n = ctypes.c_int(0)
    L.brr_hip_device_count(ctypes.byref(n))
    if n.value < 1:
        pytest.fail("no HIP device visible: GPU tests must run on an MI355X")
    return L
