import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libv2c.so")


def gpu_available() -> bool:
    try:
        import ctypes
        lib = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        if lib.hipGetDeviceCount(ctypes.byref(n)) != 0:
            return False
        return n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    from velo2cycles_amd import abi
    if not os.path.exists(abi.LIB_PATH):
        pytest.fail("libv2c.so is not built (run __graft_entry__.build())")
    return abi.load_lib()
