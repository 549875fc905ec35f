"""CPU tests of the drop-in boundary: libglsx.so loads and exports every
symbol include/glsx.h declares (no device calls), plus host-side argument
checks that fail before any device work."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        names |= set(re.findall(r"\b(glsx_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("glsx_init", "glsx_fft2d", "glsx_fft2d_dev", "glsx_phase_correlate", "glsx_buf_create",
                 "glsx_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import glscv_amd

    path = glscv_amd.lib_path()
    assert os.path.exists(path), "libglsx.so not built"
    lib = ctypes.CDLL(path)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(glsx_[a-z0-9_]+)\b", out))
    assert declared_symbols() <= exported


def test_python_signatures_cover_header():
    import glscv_amd

    assert set(glscv_amd._SIGNATURES) <= declared_symbols()


def test_null_arguments_rejected_without_device():
    import glscv_amd

    L = glscv_amd.lib()
    assert L.glsx_buf_retain(None) == -1
    assert b"null" in L.glsx_last_error()
    assert L.glsx_buf_info(None, None, None, None, None, None) == -1
    assert L.glsx_fft2d(None, None, 0, None) == -1
    assert L.glsx_phase_correlate(None, None, None, None, None) == -1
    # the chain's standalone operations check their handles before any device work
    assert L.glsx_multiply(None, None, None, None) == -1
    assert L.glsx_mul_spectrums(None, None, None, 0, None) == -1
    assert L.glsx_mul_spectrums_phase_only(None, None, None, None) == -1
    assert L.glsx_split(None, None, 0, None) == -1
    assert L.glsx_merge(None, 0, None, None) == -1
    assert L.glsx_min_max_loc(None, None, None) == -1
