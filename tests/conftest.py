import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_sessionstart(session):
    """A GPU test process holds two HIP runtimes: the one torch bundles and
    the system one libmsbg_gpu.so links. torch must take the device first;
    once the library has initialised the GPU, torch reports none. So when
    GPU tests are selected, torch initialises its device before any test."""
    if "not gpu" in (session.config.getoption("markexpr") or ""):
        return
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


@pytest.fixture(scope="session")
def have_gpu():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False
