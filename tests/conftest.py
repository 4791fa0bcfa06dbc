import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcismgpu.so")


def gpu_available() -> bool:
    try:
        from cism_amd import _lib
        L = _lib.load()
        return L.cismgpu_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    from cism_amd import _lib
    L = _lib.load()          # fails loudly when the library is missing
    assert L.cismgpu_device_count() > 0, "no HIP device visible"
    return L
