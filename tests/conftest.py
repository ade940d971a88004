import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsctl_amd.so")


def _gpu_present():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The product library, on a machine with a GPU (the test fails loudly otherwise)."""
    import sctl_amd
    if not _gpu_present():
        pytest.fail("no GPU visible: GPU tests must run on an MI355X")
    L = sctl_amd.lib()  # raises LibraryNotBuilt if the HIP library is missing
    sctl_amd.check(L.sctl_amd_init(0))
    return sctl_amd


@pytest.fixture(scope="session")
def gpu_present():
    return _gpu_present()
