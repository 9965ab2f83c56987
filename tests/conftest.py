import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs librl2hip kernels")
    config.addinivalue_line("markers", "slow: long CPU-side runs")


def gpu_available():
    try:
        import ctypes
        from radiolensfit2_amd import _native as N
        n = ctypes.c_int(0)
        N.hip().rl2_device_count(ctypes.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible (librl2hip has no CPU path)")
    return True
