"""CPU test of the server's staging copy into pinned memory
(glscv_amd/csrc/host_copy.hpp): built with g++ and run; its parts must cover
every byte once, also for byte counts that do not divide by the part count."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_staging_copy_covers_every_byte():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "test_host_copy")
        subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", os.path.join(HERE, "cpp", "test_host_copy.cpp"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "host copy ok" in r.stdout
