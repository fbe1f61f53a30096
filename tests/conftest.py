import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(TESTS, "golden")
for p in (os.path.join(ROOT, "dg-on-cuda_amd"), TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available() -> bool:
    try:
        import dgc_amd

        n = __import__("ctypes").c_int(0)
        lib = dgc_amd.lib()
        return True if lib is not None and _device_count() > 0 else False
    except Exception:
        return False


def _device_count() -> int:
    import ctypes

    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))

    return load


@pytest.fixture(scope="session")
def ctx():
    import dgc_amd

    if _device_count() == 0:
        pytest.fail("GPU test collected but no HIP device is visible")
    c = dgc_amd.Context(0)
    yield c
    c.close()
