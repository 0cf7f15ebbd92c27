import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfluxos_hip.so on the device)")


def gpu_available() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    if not oracle.LIBPATH.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def fxlib():
    from fluxos_cpp_amd import capi
    return capi.load()
