"""glscv_amd -- MI355X (gfx950) implementation of glsCV's FFT / phase-correlation path.

Python mirror of the reference's ``gls::`` API for this path, over the C ABI
of ``libglsx.so`` (``include/glsx.h``):

* ``GlsMat``          -- refcounted device buffer (glsCV/glsMat.h:45-214)
* ``fft``             -- gls::fft(src, dst, flag) (glsCV/glsFft.h:55)
* ``phaseCorrelate``  -- gls::phaseCorrelate(src1, src2, window, &response)
                         (glsCV/glsPhaseCorrelate.h:58)
* ``glsCvInit`` / ``glsCvTerminate`` (glsCV/glsCV.h:67-68)

Errors: the reference throws ``cv::Exception`` from ``GLS_Assert``; here a
non-zero status of the C ABI raises ``GlsError`` with the library's message.
There is no CPU fallback: if ``libglsx.so`` is missing or no HIP device is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Tuple

import numpy as np

__all__ = [
    "GlsError", "GlsMat", "fft", "ifft", "phaseCorrelate", "glsCvInit", "glsCvTerminate",
    "GLS_FFT_SCALE", "GLS_FFT_INVERSE", "GLS_FFT_SHIFT", "CV_32FC1", "CV_32FC2", "CV_32SC1", "CV_8UC1",
    "lib", "lib_path", "fft_dev", "phase_correlate_dev", "PocResult", "Plan", "FftDesc", "PocDesc",
    "PLAN_GRAPH", "PLAN_AUTOTUNE", "GLS_FFT_SPLIT_MFMA", "split_mfma_check", "SPLIT_MFMA_TOL",
    "magSpectrums", "logMagSpectrums", "fftMagSpectrums", "normalize", "NORM_L1", "NORM_L2", "NORM_MINMAX",
    "spectrum_dev", "normalize_dev", "fft_mag_dev", "spectrumView", "spectrum_view_dev", "HostBuffer", "stream_sync",
    "multiply", "mulSpectrums", "mulSpectrumsPhaseOnly", "split", "merge", "minMaxLoc", "MinMaxResult",
    "multiply_dev", "mul_spectrums_dev", "split_dev", "merge_dev", "min_max_loc_dev",
    "SPECTRUM_MUL", "SPECTRUM_MUL_CONJ", "SPECTRUM_PHASE_ONLY",
]

# glsCV/glsFft.h:43-47
GLS_FFT_SCALE = 2
GLS_FFT_INVERSE = 1
GLS_FFT_SHIFT = 64
# extension (no counterpart in glsFft.h): radix-16 stages on the matrix cores,
# 3-term BF16 split, honoured only for lengths that pass split_mfma_check()
GLS_FFT_SPLIT_MFMA = 0x10000
SPLIT_MFMA_TOL = 7.0710678118654752e-6   # per dimension: 1e-5 / sqrt(2)

# cv::NormTypes accepted by gls::normalize (glsCV/glsNormalize.cpp:54)
NORM_L1 = 2
NORM_L2 = 4
NORM_MINMAX = 32

CV_8UC1 = 0
CV_32SC1 = 4
CV_32FC1 = 5
CV_32FC2 = 13

_NP_OF_TYPE = {CV_8UC1: (np.uint8, 1), CV_32SC1: (np.int32, 1), CV_32FC1: (np.float32, 1),
               CV_32FC2: (np.float32, 2)}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libglsx.so")
_lib = None


class GlsError(RuntimeError):
    """Raised when a libglsx entry point returns a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"glsx error {status}: {message}")
        self.status = status


class PocResult(ctypes.Structure):
    _fields_ = [
        ("peak_x", ctypes.c_int),
        ("peak_y", ctypes.c_int),
        ("shift_x", ctypes.c_double),
        ("shift_y", ctypes.c_double),
        ("response", ctypes.c_double),
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("compute_units", ctypes.c_int),
        ("lds_per_block", ctypes.c_size_t),
        ("global_mem_bytes", ctypes.c_size_t),
    ]


class FftDesc(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("rows", ctypes.c_int), ("cols", ctypes.c_int),
                ("src_type", ctypes.c_int), ("flags", ctypes.c_int)]


class PocDesc(ctypes.Structure):
    _fields_ = [("src1", ctypes.c_void_p), ("src2", ctypes.c_void_p), ("window", ctypes.c_void_p),
                ("rows", ctypes.c_int), ("cols", ctypes.c_int)]


class PlanConfig(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("col_wcap", ctypes.c_int), ("persistent", ctypes.c_int)]


class MinMaxResult(ctypes.Structure):
    _fields_ = [("min_val", ctypes.c_float), ("max_val", ctypes.c_float), ("min_x", ctypes.c_int),
                ("min_y", ctypes.c_int), ("max_x", ctypes.c_int), ("max_y", ctypes.c_int)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("col_wcap", ctypes.c_int), ("persistent", ctypes.c_int),
                ("grid", ctypes.c_int), ("n_passes", ctypes.c_int), ("n_items", ctypes.c_int),
                ("n_levels", ctypes.c_int), ("lds_bytes", ctypes.c_size_t),
                ("uses_graph", ctypes.c_int), ("tuned", ctypes.c_int), ("tuned_us", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_int), ("max_inflight", ctypes.c_int),
                ("threads", ctypes.c_int), ("persistent", ctypes.c_int)]


class ServerCompletion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("batch_size", ctypes.c_int),
                ("result", PocResult), ("latency_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("rejected", ctypes.c_longlong),
                ("batches", ctypes.c_longlong), ("pending", ctypes.c_longlong), ("mean_batch", ctypes.c_double),
                ("budget_bytes", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t), ("inflight", ctypes.c_int),
                ("max_inflight_seen", ctypes.c_int)]


PLAN_GRAPH = 1
PLAN_AUTOTUNE = 2

# name -> (restype, argtypes); the exported surface of include/glsx.h
_VP = ctypes.c_void_p
_I = ctypes.c_int
_SIGNATURES = {
    "glsx_init": (_I, [_I]),
    "glsx_shutdown": (_I, []),
    "glsx_last_error": (ctypes.c_char_p, []),
    "glsx_get_device_info": (_I, [ctypes.POINTER(DeviceInfo)]),
    "glsx_buf_create": (_I, [_I, _I, _I, ctypes.POINTER(_VP)]),
    "glsx_buf_retain": (_I, [_VP]),
    "glsx_buf_release": (_I, [_VP]),
    "glsx_buf_release_async": (_I, [_VP, _VP]),
    "glsx_buf_info": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                           ctypes.POINTER(_VP), ctypes.POINTER(ctypes.c_size_t)]),
    "glsx_buf_upload": (_I, [_VP, _VP, ctypes.c_size_t, _VP]),
    "glsx_buf_download": (_I, [_VP, _VP, ctypes.c_size_t, _VP]),
    "glsx_buf_copy": (_I, [_VP, _VP, _VP]),
    "glsx_host_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "glsx_host_free": (_I, [_VP]),
    "glsx_buf_upload_async": (_I, [_VP, _VP, ctypes.c_size_t, _VP]),
    "glsx_buf_download_async": (_I, [_VP, _VP, ctypes.c_size_t, _VP]),
    "glsx_stream_sync": (_I, [_VP]),
    "glsx_server_submit_pinned": (_I, [_VP, _VP, _VP, _VP, _I, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "glsx_fft2d": (_I, [_VP, _VP, _I, _VP]),
    "glsx_fft2d_dev": (_I, [_VP, _I, _VP, _I, _I, _I, _VP]),
    "glsx_fft1d_dev": (_I, [_VP, _I, _VP, _I, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, _I, _VP]),
    "glsx_fft_split_mfma_check": (_I, [_I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "glsx_mag_spectrums": (_I, [_VP, _VP, _VP]),
    "glsx_log_mag_spectrums": (_I, [_VP, _VP, ctypes.c_float, _VP]),
    "glsx_spectrum_dev": (_I, [_VP, _VP, ctypes.c_longlong, _I, ctypes.c_float, _VP]),
    "glsx_normalize": (_I, [_VP, _VP, ctypes.c_double, ctypes.c_double, _I, _VP]),
    "glsx_normalize_dev": (_I, [_VP, _VP, ctypes.c_longlong, ctypes.c_double, ctypes.c_double, _I, _VP]),
    "glsx_fft2d_mag": (_I, [_VP, _VP, _I, _I, ctypes.c_float, _VP]),
    "glsx_fft2d_mag_dev": (_I, [_VP, _I, _VP, _I, _I, _I, _I, ctypes.c_float, _VP]),
    "glsx_spectrum_view": (_I, [_VP, _VP, _I, ctypes.c_float, ctypes.c_double, ctypes.c_double, _VP]),
    "glsx_spectrum_view_dev": (_I, [_VP, _I, _VP, _I, _I, _I, ctypes.c_float, ctypes.c_double, ctypes.c_double,
                                    _VP]),
    "glsx_multiply": (_I, [_VP, _VP, _VP, _VP]),
    "glsx_multiply_dev": (_I, [_VP, _VP, _VP, ctypes.c_longlong, _VP]),
    "glsx_mul_spectrums": (_I, [_VP, _VP, _VP, _I, _VP]),
    "glsx_mul_spectrums_phase_only": (_I, [_VP, _VP, _VP, _VP]),
    "glsx_mul_spectrums_dev": (_I, [_VP, _VP, _VP, ctypes.c_longlong, _I, _VP]),
    "glsx_split": (_I, [_VP, ctypes.POINTER(_VP), _I, _VP]),
    "glsx_merge": (_I, [ctypes.POINTER(_VP), _I, _VP, _VP]),
    "glsx_split_dev": (_I, [_VP, _VP, _VP, ctypes.c_longlong, _VP]),
    "glsx_merge_dev": (_I, [_VP, _VP, _VP, ctypes.c_longlong, _VP]),
    "glsx_min_max_loc": (_I, [_VP, ctypes.POINTER(MinMaxResult), _VP]),
    "glsx_min_max_loc_dev": (_I, [_VP, _I, _I, _I, _VP, _VP]),
    "glsx_plan_fft_batch": (_I, [ctypes.POINTER(FftDesc), _I, ctypes.POINTER(PlanConfig), _I,
                                 ctypes.POINTER(_VP)]),
    "glsx_poc_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(PocDesc), _I]),
    "glsx_plan_poc_batch": (_I, [ctypes.POINTER(PocDesc), _I, _VP, ctypes.c_size_t, ctypes.POINTER(PlanConfig), _I,
                                 ctypes.POINTER(_VP)]),
    "glsx_plan_execute": (_I, [_VP, _VP]),
    "glsx_plan_poc_results": (_I, [_VP, ctypes.POINTER(PocResult), _VP]),
    "glsx_plan_check": (_I, [_VP]),
    "glsx_plan_get_info": (_I, [_VP, ctypes.POINTER(PlanInfo)]),
    "glsx_plan_destroy": (_I, [_VP]),
    "glsx_server_create": (_I, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
    "glsx_server_submit": (_I, [_VP, _VP, _VP, _VP, _I, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "glsx_server_poll": (_I, [_VP, ctypes.POINTER(ServerCompletion), _I, ctypes.POINTER(_I), _I]),
    "glsx_server_get_stats": (_I, [_VP, ctypes.POINTER(ServerStats)]),
    "glsx_server_destroy": (_I, [_VP]),
    "glsx_phase_correlate": (_I, [_VP, _VP, _VP, ctypes.POINTER(PocResult), _VP]),
    "glsx_phase_correlate_dev": (_I, [_VP, _VP, _VP, _I, _I, ctypes.POINTER(PocResult), _VP]),
    "glsx_phase_correlate_dev_noalloc": (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP]),
}


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Load libglsx.so (built by ``glscv_amd.build``); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is not built; run `python -m glscv_amd.build`")
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        msg = lib().glsx_last_error()
        raise GlsError(status, msg.decode() if msg else "")


def glsCvInit(device: int = 0) -> None:
    """glsCV/glsCV.cpp:89-165 -- here: select the HIP device, build twiddle tables."""
    _check(lib().glsx_init(device))


def glsCvTerminate() -> None:
    _check(lib().glsx_shutdown())


def device_info() -> DeviceInfo:
    info = DeviceInfo()
    _check(lib().glsx_get_device_info(ctypes.byref(info)))
    return info


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)   # torch.cuda.Stream


class GlsMat:
    """Refcounted handle to a device buffer (glsCV/glsMat.h:45-214).

    ``GlsMat(ndarray)`` allocates and uploads (GlsMat(const Mat&)); ``GlsMat((rows, cols), type)``
    allocates; ``GlsMat()`` is empty.  Assignment shares the buffer; ``clone()`` copies it.
    """

    def __init__(self, arg=None, type: Optional[int] = None):
        self._h = None
        self.rows = 0
        self.cols = 0
        self._type = -1
        if arg is None:
            return
        if isinstance(arg, np.ndarray):
            arr = arg
            if arr.dtype == np.complex64:
                arr = arr.view(np.float32).reshape(arr.shape[0], arr.shape[1], 2)
            t = _type_of_array(arr)
            self._create(arr.shape[0], arr.shape[1], t)
            self.upload(arr)
        else:
            rows, cols = arg
            if type is None:
                raise GlsError(-1, "GlsMat(size, type): type required")
            self._create(int(rows), int(cols), int(type))

    def _release(self) -> None:
        if self._h is not None:
            _check(lib().glsx_buf_release(self._h))
        self._h = None
        self.rows = self.cols = 0
        self._type = -1

    def release(self, stream=None) -> None:
        """Drops this handle's reference now.  With `stream`, the storage is
        reused only after the work enqueued so far on that stream
        (glsx_buf_release_async): use it when every use of the buffer --
        including calls through its device pointer -- is ordered on `stream`."""
        if self._h is None:
            return
        if stream is None:
            _check(lib().glsx_buf_release(self._h))
        else:
            _check(lib().glsx_buf_release_async(self._h, _stream_ptr(stream)))
        self._h = None
        self.rows = self.cols = 0
        self._type = -1

    def _create(self, rows: int, cols: int, type: int) -> None:
        self._release()
        h = ctypes.c_void_p()
        _check(lib().glsx_buf_create(rows, cols, type, ctypes.byref(h)))
        self._h = h.value
        self.rows, self.cols, self._type = rows, cols, type

    def __del__(self):
        if self._h is not None and _lib is not None:
            try:
                _lib.glsx_buf_release(self._h)
            except Exception:
                pass
            self._h = None

    def _adopt(self, other: "GlsMat") -> None:
        """Make this handle refer to other's storage (dst = _dst in the reference)."""
        if other is self:
            return
        if other._h is not None:
            _check(lib().glsx_buf_retain(other._h))
        self._release()
        self._h, self.rows, self.cols, self._type = other._h, other.rows, other.cols, other._type

    def share(self) -> "GlsMat":
        """Shallow copy (GlsMat::operator=, glsMat.cpp:101-118)."""
        m = GlsMat()
        if self._h is not None:
            _check(lib().glsx_buf_retain(self._h))
            m._h, m.rows, m.cols, m._type = self._h, self.rows, self.cols, self._type
        return m

    # --- cv::Mat-like accessors (glsMat.h:128-160) ---------------------------
    def type(self) -> int:
        return self._type

    def channels(self) -> int:
        return (self._type >> 3) + 1

    def depth(self) -> int:
        return self._type & 7

    def empty(self) -> bool:
        return self._h is None

    def size(self) -> Tuple[int, int]:
        """(width, height) like cv::Size."""
        return (self.cols, self.rows)

    def total(self) -> int:
        return self.rows * self.cols

    @property
    def handle(self):
        return self._h

    @property
    def dev_ptr(self) -> int:
        p = ctypes.c_void_p()
        _check(lib().glsx_buf_info(self._h, None, None, None, ctypes.byref(p), None))
        return p.value

    def upload(self, arr: np.ndarray, stream=None) -> None:
        dt, cn = _NP_OF_TYPE[self._type]
        if arr.dtype == np.complex64:
            arr = arr.view(np.float32).reshape(arr.shape[0], arr.shape[1], 2)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape[:2] != (self.rows, self.cols) or (cn > 1 and a.shape[2] != cn):
            raise GlsError(-1, f"upload: shape {a.shape} does not match {self.rows}x{self.cols}x{cn}")
        _check(lib().glsx_buf_upload(self._h, a.ctypes.data, 0, _stream_ptr(stream)))

    def download(self, stream=None) -> np.ndarray:
        dt, cn = _NP_OF_TYPE[self._type]
        shape = (self.rows, self.cols) if cn == 1 else (self.rows, self.cols, cn)
        out = np.empty(shape, dtype=dt)
        _check(lib().glsx_buf_download(self._h, out.ctypes.data, 0, _stream_ptr(stream)))
        return out

    def download_complex(self) -> np.ndarray:
        a = self.download()
        return a.view(np.complex64).reshape(self.rows, self.cols)

    def _check_host(self, hb: "HostBuffer", what: str) -> None:
        if hb.nbytes != self.rows * self.cols * _NP_OF_TYPE[self._type][1] * np.dtype(_NP_OF_TYPE[self._type][0]).itemsize:
            raise GlsError(-1, f"{what}: host buffer of {hb.nbytes} bytes does not match {self.rows}x{self.cols}")

    def upload_async(self, hb: "HostBuffer", stream=None) -> None:
        """Stream-ordered upload from pinned memory; hb must stay unchanged until the stream passes it."""
        self._check_host(hb, "upload_async")
        _check(lib().glsx_buf_upload_async(self._h, hb.ptr, 0, _stream_ptr(stream)))

    def download_async(self, hb: "HostBuffer", stream=None) -> None:
        """Stream-ordered download into pinned memory; valid after the stream is synchronised."""
        self._check_host(hb, "download_async")
        _check(lib().glsx_buf_download_async(self._h, hb.ptr, 0, _stream_ptr(stream)))

    def copyTo(self, m: "GlsMat") -> "GlsMat":
        if m.empty() or m.rows != self.rows or m.cols != self.cols or m._type != self._type:
            m._create(self.rows, self.cols, self._type)
        _check(lib().glsx_buf_copy(self._h, m._h, None))
        return m

    def clone(self) -> "GlsMat":
        m = GlsMat((self.rows, self.cols), self._type)
        _check(lib().glsx_buf_copy(self._h, m._h, None))
        return m


def _type_of_array(arr: np.ndarray) -> int:
    if arr.ndim == 2:
        cn = 1
    elif arr.ndim == 3:
        cn = arr.shape[2]
    else:
        raise GlsError(-1, "GlsMat: 2D (or 2D x channels) array expected")
    for t, (dt, c) in _NP_OF_TYPE.items():
        if np.dtype(dt) == arr.dtype and c == cn:
            return t
    raise GlsError(-1, f"GlsMat: unsupported dtype/channels {arr.dtype}/{cn}")


def fft(src: GlsMat, dst: Optional[GlsMat] = None, flag: int = 0, stream=None) -> GlsMat:
    """gls::fft (glsCV/glsFft.cpp:675-708).  Returns dst (reused when it is a
    CV_32FC2 buffer of the source size, otherwise newly allocated; in place
    fft(C, C) is allowed)."""
    if src.empty():
        raise GlsError(-1, "fft: empty source")
    if dst is not None and not (dst.empty() or dst.rows != src.rows or dst.cols != src.cols
                                or dst.type() != CV_32FC2):
        _check(lib().glsx_fft2d(src.handle, dst.handle, int(flag), _stream_ptr(stream)))
        return dst
    # getDstMat (glsMat.h:173-214): a fresh buffer, then dst refers to it (dst
    # may be the source handle itself, e.g. fft(frame, frame) on a real frame)
    fresh = GlsMat((src.rows, src.cols), CV_32FC2)
    _check(lib().glsx_fft2d(src.handle, fresh.handle, int(flag), _stream_ptr(stream)))
    if dst is None:
        return fresh
    dst._adopt(fresh)
    return dst


def ifft(src: GlsMat, dst: Optional[GlsMat] = None, scale: bool = True, stream=None) -> GlsMat:
    """Sugar for fft(src, dst, GLS_FFT_INVERSE [| GLS_FFT_SCALE])."""
    return fft(src, dst, GLS_FFT_INVERSE | (GLS_FFT_SCALE if scale else 0), stream)


def phaseCorrelate(src1: GlsMat, src2: GlsMat, window: Optional[GlsMat] = None, stream=None):
    """gls::phaseCorrelate (glsCV/glsPhaseCorrelate.cpp:142-236).

    Returns ((shift_x, shift_y), response, (peak_x, peak_y)).
    """
    res = PocResult()
    win = window.handle if (window is not None and not window.empty()) else None
    _check(lib().glsx_phase_correlate(src1.handle, src2.handle, win, ctypes.byref(res), _stream_ptr(stream)))
    return (res.shift_x, res.shift_y), res.response, (res.peak_x, res.peak_y)


def phase_correlate_dev_noalloc(src1_ptr: int, src2_ptr: int, window_ptr: Optional[int], rows: int, cols: int,
                                dev_out_ptr: int, stream=None) -> None:
    """glsx_phase_correlate_dev_noalloc: the same chain on the library's static
    workspace; writes the 32-byte result to device memory at dev_out_ptr in
    stream order (no allocation, no synchronisation)."""
    _check(lib().glsx_phase_correlate_dev_noalloc(src1_ptr, src2_ptr, window_ptr, rows, cols, dev_out_ptr,
                                                  _stream_ptr(stream)))


# --- raw device-pointer entry points (torch tensors as plumbing) -------------
def fft_dev(src_ptr: int, src_type: int, dst_ptr: int, rows: int, cols: int, flags: int = 0, stream=None) -> None:
    _check(lib().glsx_fft2d_dev(src_ptr, src_type, dst_ptr, rows, cols, flags, _stream_ptr(stream)))


def fft1d_dev(src_ptr: int, src_type: int, dst_ptr: int, length: int, count: int, elem_stride: int,
              batch_stride: int, flags: int = 0, stream=None) -> None:
    _check(lib().glsx_fft1d_dev(src_ptr, src_type, dst_ptr, length, count, elem_stride, batch_stride, flags,
                                _stream_ptr(stream)))


# --- spectrum display chain (glsCam/glsCam.cpp:196-202) ----------------------
def _plane_for(src: GlsMat, dst: Optional[GlsMat], may_alias: bool) -> GlsMat:
    """getDstMat for a CV_32FC1 plane of src's size."""
    if (dst is not None and not dst.empty() and dst.rows == src.rows and dst.cols == src.cols
            and dst.type() == CV_32FC1 and (may_alias or dst.handle != src.handle)):
        return dst
    return GlsMat((src.rows, src.cols), CV_32FC1)


def _assign(dst: Optional[GlsMat], out: GlsMat) -> GlsMat:
    if dst is None or out is dst:
        return out
    dst._adopt(out)
    return dst


def magSpectrums(src: GlsMat, dst: Optional[GlsMat] = None, stream=None) -> GlsMat:
    """gls::magSpectrums (glsBasicOperation.cpp:337-343): |z| of a CV_32FC2 spectrum."""
    out = _plane_for(src, dst, False)
    _check(lib().glsx_mag_spectrums(src.handle, out.handle, _stream_ptr(stream)))
    return _assign(dst, out)


def logMagSpectrums(src: GlsMat, dst: Optional[GlsMat] = None, offset: float = 1.0, stream=None) -> GlsMat:
    """gls::logMagSpectrums (glsBasicOperation.cpp:345-351): log(|z| + offset)."""
    out = _plane_for(src, dst, False)
    _check(lib().glsx_log_mag_spectrums(src.handle, out.handle, float(offset), _stream_ptr(stream)))
    return _assign(dst, out)


def fftMagSpectrums(src: GlsMat, dst: Optional[GlsMat] = None, flag: int = 0, log_mag: bool = True,
                    offset: float = 1.0, stream=None) -> GlsMat:
    """fft(src, flag) then (log)magSpectrums, fused into the FFT's last pass."""
    out = _plane_for(src, dst, False)
    _check(lib().glsx_fft2d_mag(src.handle, out.handle, int(flag), 1 if log_mag else 0, float(offset),
                                _stream_ptr(stream)))
    return _assign(dst, out)


def spectrumView(src: GlsMat, dst: Optional[GlsMat] = None, flag: int = GLS_FFT_SHIFT, offset: float = 1.0,
                 alpha: float = 0.0, beta: float = 1.0, stream=None) -> GlsMat:
    """glsCam's FFT view: fft(flag) -> logMagSpectrums(offset) ->
    normalize(alpha, beta, NORM_MINMAX), in three launches."""
    out = _plane_for(src, dst, False)
    _check(lib().glsx_spectrum_view(src.handle, out.handle, int(flag), float(offset), float(alpha), float(beta),
                                    _stream_ptr(stream)))
    return _assign(dst, out)


def spectrum_view_dev(src_ptr: int, src_type: int, dst_ptr: int, rows: int, cols: int, flags: int = GLS_FFT_SHIFT,
                      offset: float = 1.0, alpha: float = 0.0, beta: float = 1.0, stream=None):
    _check(lib().glsx_spectrum_view_dev(src_ptr, src_type, dst_ptr, rows, cols, int(flags), float(offset),
                                        float(alpha), float(beta), _stream_ptr(stream)))


def normalize(src: GlsMat, dst: Optional[GlsMat] = None, alpha: float = 1.0, beta: float = 0.0,
              normType: int = NORM_L2, stream=None) -> GlsMat:
    """gls::normalize (glsNormalize.cpp:52-75) on a CV_32FC1 plane; in place allowed."""
    out = _plane_for(src, dst, True)
    _check(lib().glsx_normalize(src.handle, out.handle, float(alpha), float(beta), int(normType),
                                _stream_ptr(stream)))
    return _assign(dst, out)


# --- the chain's operations one at a time (glsPhaseCorrelate.cpp:174-210) -------
SPECTRUM_MUL = 0
SPECTRUM_MUL_CONJ = 1
SPECTRUM_PHASE_ONLY = 2


def _dst_for(src0: GlsMat, src1: GlsMat, dst: Optional[GlsMat]) -> GlsMat:
    """getDstMat(src0, src1, dst) (glsMat.h:185-196): a new image when dst is
    one of the sources or differs in size or type."""
    if (dst is not None and not dst.empty() and dst.rows == src0.rows and dst.cols == src0.cols
            and dst.type() == src0.type() and dst.handle not in (src0.handle, src1.handle)):
        return dst
    return GlsMat((src0.rows, src0.cols), src0.type())


def multiply(src0: GlsMat, src1: GlsMat, dst: Optional[GlsMat] = None, stream=None) -> GlsMat:
    """gls::multiply (glsBasicOperation.cpp:265-272): per-element product."""
    out = _dst_for(src0, src1, dst)
    _check(lib().glsx_multiply(src0.handle, src1.handle, out.handle, _stream_ptr(stream)))
    return _assign(dst, out)


def mulSpectrums(src0: GlsMat, src1: GlsMat, dst: Optional[GlsMat] = None, conj: bool = False,
                 stream=None) -> GlsMat:
    """gls::mulSpectrums (glsBasicOperation.cpp:320-327): src0*src1 or src0*conj(src1)."""
    out = _dst_for(src0, src1, dst)
    _check(lib().glsx_mul_spectrums(src0.handle, src1.handle, out.handle, 1 if conj else 0, _stream_ptr(stream)))
    return _assign(dst, out)


def mulSpectrumsPhaseOnly(src0: GlsMat, src1: GlsMat, dst: Optional[GlsMat] = None, stream=None) -> GlsMat:
    """gls::mulSpectrumsPhaseOnly (glsBasicOperation.cpp:329-335): c/|c|, c = src0*conj(src1)."""
    out = _dst_for(src0, src1, dst)
    _check(lib().glsx_mul_spectrums_phase_only(src0.handle, src1.handle, out.handle, _stream_ptr(stream)))
    return _assign(dst, out)


def split(src: GlsMat, stream=None) -> List[GlsMat]:
    """gls::split (glsSplit.cpp:192-210): one new CV_32FC1 image per channel."""
    planes = [GlsMat((src.rows, src.cols), CV_32FC1) for _ in range(src.channels())]
    arr = (_VP * len(planes))(*[p.handle for p in planes])
    _check(lib().glsx_split(src.handle, arr, len(planes), _stream_ptr(stream)))
    return planes


def merge(planes: List[GlsMat], dst: Optional[GlsMat] = None, stream=None) -> GlsMat:
    """gls::merge (glsMerge.cpp:193-211): 1 or 2 CV_32FC1 planes into one image."""
    t = CV_32FC1 if len(planes) == 1 else CV_32FC2
    p0 = planes[0]
    if (dst is not None and not dst.empty() and dst.rows == p0.rows and dst.cols == p0.cols
            and dst.type() == t):
        out = dst
    else:
        out = GlsMat((p0.rows, p0.cols), t)
    arr = (_VP * len(planes))(*[p.handle for p in planes])
    _check(lib().glsx_merge(arr, len(planes), out.handle, _stream_ptr(stream)))
    return _assign(dst, out)


def minMaxLoc(src: GlsMat, stream=None):
    """gls::minMaxLoc (glsMinMaxLoc.cpp:149-199): (minVal, maxVal, (min_x, min_y),
    (max_x, max_y)) of channel 0, first extrema in row-major order."""
    r = MinMaxResult()
    _check(lib().glsx_min_max_loc(src.handle, ctypes.byref(r), _stream_ptr(stream)))
    return float(r.min_val), float(r.max_val), (r.min_x, r.min_y), (r.max_x, r.max_y)


def multiply_dev(a_ptr: int, b_ptr: int, dst_ptr: int, n: int, stream=None) -> None:
    _check(lib().glsx_multiply_dev(a_ptr, b_ptr, dst_ptr, n, _stream_ptr(stream)))


def mul_spectrums_dev(a_ptr: int, b_ptr: int, dst_ptr: int, n: int, mode: int, stream=None) -> None:
    _check(lib().glsx_mul_spectrums_dev(a_ptr, b_ptr, dst_ptr, n, int(mode), _stream_ptr(stream)))


def split_dev(src_ptr: int, plane0_ptr: int, plane1_ptr: Optional[int], n: int, stream=None) -> None:
    _check(lib().glsx_split_dev(src_ptr, plane0_ptr, plane1_ptr, n, _stream_ptr(stream)))


def merge_dev(plane0_ptr: int, plane1_ptr: int, dst_ptr: int, n: int, stream=None) -> None:
    _check(lib().glsx_merge_dev(plane0_ptr, plane1_ptr, dst_ptr, n, _stream_ptr(stream)))


def min_max_loc_dev(src_ptr: int, src_type: int, rows: int, cols: int, dev_out_ptr: int, stream=None) -> None:
    """Result (a MinMaxResult, 24 bytes) written to device memory in stream order."""
    _check(lib().glsx_min_max_loc_dev(src_ptr, int(src_type), rows, cols, dev_out_ptr, _stream_ptr(stream)))


def spectrum_dev(src_ptr: int, dst_ptr: int, n: int, log_mag: bool = False, offset: float = 0.0, stream=None):
    _check(lib().glsx_spectrum_dev(src_ptr, dst_ptr, n, 1 if log_mag else 0, float(offset), _stream_ptr(stream)))


def normalize_dev(src_ptr: int, dst_ptr: int, n: int, alpha: float, beta: float, norm_type: int, stream=None):
    _check(lib().glsx_normalize_dev(src_ptr, dst_ptr, n, float(alpha), float(beta), int(norm_type),
                                    _stream_ptr(stream)))


def fft_mag_dev(src_ptr: int, src_type: int, dst_ptr: int, rows: int, cols: int, flags: int = 0,
                log_mag: bool = True, offset: float = 1.0, stream=None):
    _check(lib().glsx_fft2d_mag_dev(src_ptr, src_type, dst_ptr, rows, cols, flags, 1 if log_mag else 0,
                                    float(offset), _stream_ptr(stream)))


def split_mfma_check(length: int) -> Tuple[float, bool]:
    """FP64 check of the split-precision MFMA pass for one transform length.

    Returns (relative RMS error vs a direct FP64 DFT, enabled): the flag
    GLS_FFT_SPLIT_MFMA is honoured for this length only when enabled."""
    err = ctypes.c_double(0.0)
    en = ctypes.c_int(0)
    _check(lib().glsx_fft_split_mfma_check(int(length), ctypes.byref(err), ctypes.byref(en)))
    return err.value, bool(en.value)


class Plan:
    """A batched plan: many FFTs (or phase correlations) in one persistent launch."""

    def __init__(self, handle, kind, keep=None):
        self._h = handle
        self.kind = kind
        self._keep = keep            # objects whose device memory the plan points at
        self.count = 0

    @classmethod
    def fft_batch(cls, descs, options: int = 0, keep=None, threads: int = 0, col_wcap: int = 16,
                  persistent: int = 0) -> "Plan":
        arr = (FftDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        cfg = ctypes.byref(PlanConfig(threads, col_wcap, persistent)) if threads else None
        _check(lib().glsx_plan_fft_batch(arr, len(descs), cfg, options, ctypes.byref(h)))
        p = cls(h.value, "fft", keep)
        p.count = len(descs)
        return p

    @classmethod
    def poc_batch(cls, descs, workspace_ptr: int, workspace_bytes: int, threads: int = 0, options: int = 0,
                  keep=None, col_wcap: int = 16, persistent: int = 0) -> "Plan":
        arr = (PocDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        cfg = ctypes.byref(PlanConfig(threads, col_wcap, persistent)) if threads else None
        _check(lib().glsx_plan_poc_batch(arr, len(descs), workspace_ptr, workspace_bytes, cfg, options,
                                         ctypes.byref(h)))
        p = cls(h.value, "poc", keep)
        p.count = len(descs)
        return p

    @staticmethod
    def poc_workspace_bytes(descs) -> int:
        arr = (PocDesc * len(descs))(*descs)
        return int(lib().glsx_poc_batch_workspace_bytes(arr, len(descs)))

    def execute(self, stream=None) -> None:
        _check(lib().glsx_plan_execute(self._h, _stream_ptr(stream)))

    def check(self) -> None:
        _check(lib().glsx_plan_check(self._h))

    def results(self, stream=None):
        out = (PocResult * self.count)()
        _check(lib().glsx_plan_poc_results(self._h, out, _stream_ptr(stream)))
        return list(out)

    def info(self) -> PlanInfo:
        i = PlanInfo()
        _check(lib().glsx_plan_get_info(self._h, ctypes.byref(i)))
        return i

    def destroy(self) -> None:
        if self._h is not None:
            _check(lib().glsx_plan_destroy(self._h))
            self._h = None

    def __del__(self):
        if self._h is not None and _lib is not None:
            try:
                _lib.glsx_plan_destroy(self._h)
            except Exception:
                pass
            self._h = None


class HostBuffer:
    """Pinned host memory (glsx_host_alloc) viewed as a numpy array: the
    source / destination of asynchronous GlsMat transfers and of zero-copy
    server submissions."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(d) for d in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().glsx_host_alloc(max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value
        raw = (ctypes.c_byte * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(raw, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self.array = None
            _check(lib().glsx_host_free(self.ptr))
            self.ptr = None

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            try:
                _lib.glsx_host_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


def stream_sync(stream=None) -> None:
    _check(lib().glsx_stream_sync(_stream_ptr(stream)))


class PocServer:
    """Continuous-batching phase-correlation server (include/glsx.h glsx_server_*)."""

    def __init__(self, hbm_budget_bytes: int = 8 << 30, max_batch: int = 16, max_inflight: int = 3,
                 threads: int = 0, persistent: int = 2):
        h = ctypes.c_void_p()
        cfg = ServerConfig(hbm_budget_bytes, max_batch, max_inflight, threads, persistent)
        _check(lib().glsx_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self._held = {}   # request id -> pinned buffers the server still reads

    def submit_pinned(self, img1: HostBuffer, img2: HostBuffer, window: Optional[HostBuffer] = None) -> int:
        """Zero-copy submission of images in pinned host memory; the buffers are
        held until the request's completion is polled."""
        rows, cols = img1.shape
        rid = ctypes.c_longlong()
        _check(lib().glsx_server_submit_pinned(self._h, img1.ptr, img2.ptr, window.ptr if window else None,
                                               rows, cols, ctypes.byref(rid)))
        self._held[rid.value] = (img1, img2, window)
        return rid.value

    def submit(self, img1: np.ndarray, img2: np.ndarray, window: Optional[np.ndarray] = None) -> int:
        a = np.ascontiguousarray(img1, dtype=np.float32)
        b = np.ascontiguousarray(img2, dtype=np.float32)
        w = np.ascontiguousarray(window, dtype=np.float32) if window is not None else None
        rid = ctypes.c_longlong()
        _check(lib().glsx_server_submit(self._h, a.ctypes.data, b.ctypes.data, w.ctypes.data if w is not None else None,
                                        a.shape[0], a.shape[1], ctypes.byref(rid)))
        return rid.value

    def poll(self, max_items: int = 256, timeout_ms: int = 0):
        buf = (ServerCompletion * max_items)()
        n = ctypes.c_int()
        _check(lib().glsx_server_poll(self._h, buf, max_items, ctypes.byref(n), timeout_ms))
        out = [buf[i] for i in range(n.value)]
        for c in out:
            self._held.pop(c.id, None)
        return out

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().glsx_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def close(self) -> None:
        if self._h is not None:
            _check(lib().glsx_server_destroy(self._h))
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            try:
                _lib.glsx_server_destroy(self._h)
            except Exception:
                pass
            self._h = None


def phase_correlate_dev(a_ptr: int, b_ptr: int, win_ptr: Optional[int], rows: int, cols: int,
                        stream=None) -> PocResult:
    res = PocResult()
    _check(lib().glsx_phase_correlate_dev(a_ptr, b_ptr, win_ptr, rows, cols, ctypes.byref(res),
                                          _stream_ptr(stream)))
    return res
