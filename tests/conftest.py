import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def gpu():
    """The loaded HIP library; GPU tests fail loudly when it is missing or no device is usable.

    torch (used by some tests for device buffers) brings its own HIP runtime
    library; it is initialised first so that the HIP library here binds to
    the same runtime instead of loading a second one."""
    try:
        import torch
        torch.cuda.init()
    except ImportError:
        pass
    import uaamg_amd
    lib = uaamg_amd.load()
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.L.uaamg_create(ctypes.byref(h), None)
    assert rc == 0, f"no usable gfx950 device (uaamg_create -> {uaamg_amd.status_name(rc)})"
    lib.L.uaamg_destroy(h)
    return lib
