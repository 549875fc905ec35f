"""Runs the C++ gls:: API test program (tests/cpp/test_gls_api.cpp, built by
glscv_amd.build next to libglsx.so) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_gls_api")


def test_cpp_gls_api(gls):
    assert os.path.exists(BIN), "tests/cpp/test_gls_api not built (run the build entry point)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
