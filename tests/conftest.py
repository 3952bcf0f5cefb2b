import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'oracle'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: long-running')


def gpu_available():
    """Count HIP devices without importing torch."""
    import ctypes
    try:
        hip = ctypes.CDLL('libamdhip64.so')
    except OSError:
        try:
            hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
        except OSError:
            return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


@pytest.fixture(scope='session')
def gpu():
    if not gpu_available():
        pytest.fail('GPU test selected but no GPU is visible')
    return 0
