import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available() -> bool:
    try:
        from cfd_nssolver_forhpc_amd import capi
        import ctypes as C
        n = C.c_int()
        capi.load().cfd_device_count(C.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    from cfd_nssolver_forhpc_amd import capi
    if not capi.LIB_PATH.exists():
        pytest.fail("HIP library not built: the GPU tests need cfd_nssolver_forhpc_amd/_lib/libcfd_hip.so")
    if not gpu_available():
        pytest.fail("no HIP device visible")
    return True
