import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


_GPU = None


def gpu_available() -> bool:
    """True when the HIP runtime reports a device (asked once, via ctypes)."""
    global _GPU
    if _GPU is None:
        import ctypes

        try:
            hip = ctypes.CDLL("libamdhip64.so")
            cnt = ctypes.c_int(0)
            _GPU = hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0
        except OSError:
            _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def stellar():
    import stellarsolver_amd as S

    S.load_library()
    return S
