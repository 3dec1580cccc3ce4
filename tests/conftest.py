import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libraries")


def gpu_available() -> bool:
    try:
        from annarchy_amd import _lib
        import ctypes
        n = ctypes.c_int32(0)
        rc = _lib.lib().ahip_device_count(ctypes.byref(n))
        return rc == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device / library is available")
    return True
