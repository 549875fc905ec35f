import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libglsx.so")


@pytest.fixture(scope="session")
def gls():
    """The product library, initialised on device 0.  Raises (does not skip)
    when libglsx.so is missing; skips only when no HIP device exists."""
    import glscv_amd

    glscv_amd.lib()
    try:
        glscv_amd.glsCvInit(0)
    except glscv_amd.GlsError as e:
        if "not present" in str(e) or "no ROCm-capable device" in str(e) or "hipGetDeviceCount" in str(e):
            pytest.skip(f"no HIP device: {e}")
        raise
    return glscv_amd
