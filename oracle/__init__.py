"""CPU oracle for glsCV's FFT / phase-correlation path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py`` -- always as the checker or as the
timed CPU baseline, never as the thing measured or shipped.  The product
library (``glscv_amd``) does not import this package.

The C sources (``oracle.c``) restate, with file:line citations:
  * ``UnitTest_glsCV/fft_stockhum.cpp:771-1033`` -- FP32 radix-4 Stockham
    (``fft_r4_f32``), with the flag handling of ``glsCV/glsFft.cpp:586-664``;
  * ``glsCV/glsBasicOperation.cpp:195-206`` -- POC (``poc``);
  * ``glsCV/glsBasicOperation.cpp:320-327`` -- spectrum products (``mul_spectrums``);
  * ``glsCV/glsMinMaxLoc.cpp:65-118`` -- first-maximum argmax (``argmax_first``),
    and with ``:149-199`` both passes of ``gls::minMaxLoc`` (``min_max_loc``);
  * ``glsCV/glsPhaseCorrelate.cpp:60-231`` -- centroid and chain
    (``weighted_centroid``, ``phase_correlate``);
  * ``glsCV/glsBasicOperation.cpp:337-351`` and ``glsCV/glsNormalize.cpp:52-75``
    -- the spectrum display chain (``mag_spectrums``, ``log_mag_spectrums``,
    ``normalize``);
plus an FP64 Stockham radix-2 FFT (``fft_f64``) used as the numerical ground
truth.  Every function is pinned by tests in ``tests/test_oracle.py``
(brute-force DFT, numpy golden vectors, closed forms).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

INVERSE = 1
SCALE = 2
SHIFT = 64

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class PocResult(ctypes.Structure):
    _fields_ = [
        ("peak_x", ctypes.c_int),
        ("peak_y", ctypes.c_int),
        ("shift_x", ctypes.c_double),
        ("shift_y", ctypes.c_double),
        ("response", ctypes.c_double),
    ]


def build() -> str:
    """Compile liboracle.so from oracle.c with the Makefile beside it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_fft2d_f64.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fft2d_r4_f32.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fft2d_r4_f32_mt.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_plant_fault.argtypes = [ctypes.c_int]
        L.orc_plant_fault.restype = None
        L.orc_poc_f32.argtypes = [fp, fp, fp, ctypes.c_long]
        L.orc_poc_f32.restype = None
        L.orc_mul_spectrums_f32.argtypes = [fp, fp, fp, ctypes.c_long, ctypes.c_int]
        L.orc_mul_spectrums_f32.restype = None
        L.orc_min_max_loc_f32.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, fp, ip, ip, ip, ip]
        L.orc_argmax_first_f32.argtypes = [fp, ctypes.c_int, ctypes.c_int, ip, ip, fp]
        L.orc_argmax_first_f32.restype = None
        L.orc_weighted_centroid.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_weighted_centroid.restype = None
        L.orc_phase_correlate.argtypes = [fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(PocResult), fp]
        L.orc_mag_spectrums_f32.argtypes = [fp, fp, ctypes.c_long]
        L.orc_mag_spectrums_f32.restype = None
        L.orc_log_mag_spectrums_f32.argtypes = [fp, fp, ctypes.c_long, ctypes.c_float]
        L.orc_log_mag_spectrums_f32.restype = None
        L.orc_normalize_f32.argtypes = [fp, fp, ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def plant_fault(which: int) -> None:
    """Test hook: plant a known mistake in the oracle (0 removes it)."""
    lib().orc_plant_fault(int(which))


def fft_f64(x: np.ndarray, flags: int = 0) -> np.ndarray:
    """FP64 2D DFT of a complex (rows, cols) array with gls::fft flags."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    if x.ndim == 1:
        x = x.reshape(1, -1)
    rows, cols = x.shape
    out = np.empty_like(x)
    rc = lib().orc_fft2d_f64(_ptr(x.view(np.float64), ctypes.c_double),
                             _ptr(out.view(np.float64), ctypes.c_double), rows, cols, flags)
    if rc:
        raise ValueError(f"orc_fft2d_f64 failed ({rc}) for {rows}x{cols}")
    return out


def fft_r4_f32(x: np.ndarray, flags: int = 0, nthreads: int = 1) -> np.ndarray:
    """FP32 restated radix-4 Stockham (fft_stockhum.cpp:771-1033) on complex64."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex64))
    if x.ndim == 1:
        x = x.reshape(1, -1)
    rows, cols = x.shape
    out = np.empty_like(x)
    rc = lib().orc_fft2d_r4_f32_mt(_ptr(x.view(np.float32), ctypes.c_float),
                                   _ptr(out.view(np.float32), ctypes.c_float), rows, cols, flags, nthreads)
    if rc:
        raise ValueError(f"orc_fft2d_r4_f32 failed ({rc}) for {rows}x{cols}")
    return out


def poc(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex64)
    b = np.ascontiguousarray(b, dtype=np.complex64)
    c = np.empty_like(a)
    lib().orc_poc_f32(_ptr(a.view(np.float32), ctypes.c_float), _ptr(b.view(np.float32), ctypes.c_float),
                      _ptr(c.view(np.float32), ctypes.c_float), a.size)
    return c


def mul_spectrums(a: np.ndarray, b: np.ndarray, conj: bool = False) -> np.ndarray:
    """Restated gls::mulSpectrums on complex64: a*b, or a*conj(b)."""
    a = np.ascontiguousarray(a, dtype=np.complex64)
    b = np.ascontiguousarray(b, dtype=np.complex64)
    c = np.empty_like(a)
    lib().orc_mul_spectrums_f32(_ptr(a.view(np.float32), ctypes.c_float), _ptr(b.view(np.float32), ctypes.c_float),
                                _ptr(c.view(np.float32), ctypes.c_float), a.size, 1 if conj else 0)
    return c


def min_max_loc(v: np.ndarray):
    """Restated gls::minMaxLoc of channel 0 of a (rows, cols) or (rows, cols, cn)
    float32 image: (min_val, max_val, (min_x, min_y), (max_x, max_y))."""
    v = _f32(v)
    rows, cols = v.shape[:2]
    cn = v.shape[2] if v.ndim == 3 else 1
    mn, mx = ctypes.c_float(), ctypes.c_float()
    a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib().orc_min_max_loc_f32(_ptr(v, ctypes.c_float), rows, cols, cn, ctypes.byref(mn), ctypes.byref(mx),
                                   ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
    if rc:
        raise ValueError(f"orc_min_max_loc_f32 failed ({rc})")
    return mn.value, mx.value, (a.value, b.value), (c.value, d.value)


def argmax_first(v: np.ndarray):
    v = _f32(v)
    rows, cols = v.shape
    px, py, pm = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
    lib().orc_argmax_first_f32(_ptr(v, ctypes.c_float), rows, cols, ctypes.byref(px), ctypes.byref(py),
                               ctypes.byref(pm))
    return px.value, py.value, pm.value


def weighted_centroid(v: np.ndarray, peak_x: int, peak_y: int, box=(5, 5)):
    v = _f32(v)
    rows, cols = v.shape
    cx, cy, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_weighted_centroid(_ptr(v, ctypes.c_float), rows, cols, peak_x, peak_y, box[0], box[1],
                                ctypes.byref(cx), ctypes.byref(cy), ctypes.byref(r))
    return cx.value, cy.value, r.value


def phase_correlate(img1, img2, window=None, nthreads: int = 1, want_surface: bool = False):
    """Restated gls::phaseCorrelate chain; returns (PocResult, surface or None)."""
    a = _f32(img1)
    b = _f32(img2)
    rows, cols = a.shape
    w = _f32(window) if window is not None else None
    res = PocResult()
    surf = np.empty((rows, cols), np.float32) if want_surface else None
    rc = lib().orc_phase_correlate(_ptr(a, ctypes.c_float), _ptr(b, ctypes.c_float),
                                   _ptr(w, ctypes.c_float) if w is not None else None, rows, cols, nthreads,
                                   ctypes.byref(res),
                                   _ptr(surf, ctypes.c_float) if surf is not None else None)
    if rc:
        raise ValueError(f"orc_phase_correlate failed ({rc})")
    return res, surf


NORM_L1 = 2
NORM_L2 = 4
NORM_MINMAX = 32


def mag_spectrums(c: np.ndarray) -> np.ndarray:
    """Restated gls::magSpectrums: |z| of a complex plane, float32."""
    c = np.ascontiguousarray(c, dtype=np.complex64)
    m = np.empty(c.shape, np.float32)
    lib().orc_mag_spectrums_f32(_ptr(c.view(np.float32), ctypes.c_float), _ptr(m, ctypes.c_float), c.size)
    return m


def log_mag_spectrums(c: np.ndarray, offset: float = 1.0) -> np.ndarray:
    """Restated gls::logMagSpectrums: log(|sqrt(re^2 + im^2) + offset|)."""
    c = np.ascontiguousarray(c, dtype=np.complex64)
    m = np.empty(c.shape, np.float32)
    lib().orc_log_mag_spectrums_f32(_ptr(c.view(np.float32), ctypes.c_float), _ptr(m, ctypes.c_float), c.size,
                                    float(offset))
    return m


def normalize(src: np.ndarray, alpha: float = 1.0, beta: float = 0.0, norm_type: int = NORM_L2) -> np.ndarray:
    """Restated gls::normalize (NORM_MINMAX, NORM_L1, NORM_L2) on a float32 plane."""
    s = _f32(src)
    d = np.empty_like(s)
    rc = lib().orc_normalize_f32(_ptr(s, ctypes.c_float), _ptr(d, ctypes.c_float), s.size, float(alpha),
                                 float(beta), int(norm_type))
    if rc:
        raise ValueError(f"orc_normalize_f32 failed ({rc})")
    return d


def rel_rms(y, ref) -> float:
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((y - ref).ravel())
    return float(num / den) if den > 0 else float(num)
