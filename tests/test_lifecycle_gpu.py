"""Life cycle of the library's device state and GlsMat storage.

* glsx_shutdown refuses (GLSX_ERR_STATE) while a plan or a server is alive:
  both hold the device's twiddle tables (the reference's glsCvTerminate,
  glsCV.cpp, has no such objects).  Nothing stale is ever executed here.
* The last release of a GlsMat is stream-ordered (glsMat.cpp:121-125 frees the
  texture at once): it does not wait for the device -- work on another stream
  is still running when it returns -- and its block is not handed to a new
  buffer before the work that used it has finished.
"""
import ctypes
import gc

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_ERROR_NOT_READY = 600


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.fail("libamdhip64.so not found")


def _nonblocking_stream(hip):
    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), ctypes.c_uint(1)) == 0  # hipStreamNonBlocking
    return s.value


def _destroy(hip, s):
    hip.hipStreamSynchronize(ctypes.c_void_p(s))
    hip.hipStreamDestroy(ctypes.c_void_p(s))


def _rand_c64(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)


def test_shutdown_refused_while_a_plan_is_alive(gls):
    gc.collect()  # plans of earlier tests that are no longer referenced
    x = _rand_c64(64, 1)
    a, b = gls.GlsMat(x), gls.GlsMat((64, 64), gls.CV_32FC2)
    plan = gls.Plan.fft_batch([gls.FftDesc(a.dev_ptr, b.dev_ptr, 64, 64, gls.CV_32FC2, 0)], 0, keep=(a, b))
    with pytest.raises(gls.GlsError) as e:
        gls.glsCvTerminate()
    assert e.value.status == -5 and "alive" in str(e.value)
    # the refusal changed nothing: the plan still runs and is right
    plan.execute()
    plan.check()
    y = b.download_complex()
    assert np.linalg.norm(y - np.fft.fft2(x)) / np.linalg.norm(np.fft.fft2(x)) < 1e-5
    plan.destroy()
    gls.glsCvTerminate()  # allowed once the plan is gone
    gls.glsCvInit(0)
    # buffers that outlived the shutdown still work
    z = gls.fft(a).download_complex()
    assert np.array_equal(z.view(np.uint32), gls.fft(gls.GlsMat(x)).download_complex().view(np.uint32))


def test_shutdown_refused_while_a_server_is_alive(gls):
    gc.collect()
    srv = gls.PocServer(hbm_budget_bytes=64 << 20, max_batch=4, max_inflight=1)
    try:
        with pytest.raises(gls.GlsError) as e:
            gls.glsCvTerminate()
        assert e.value.status == -5
        rng = np.random.default_rng(3)
        a = rng.uniform(0, 1, (64, 64)).astype(np.float32)
        b = np.roll(a, (5, -7), axis=(0, 1)).astype(np.float32)
        rid = srv.submit(a, b)
        got = []
        for _ in range(100):
            got += srv.poll(8, timeout_ms=100)
            if got:
                break
        assert got and got[0].id == rid and got[0].status == 0
        assert (got[0].result.peak_x, got[0].result.peak_y) == (32 + 7, 32 - 5)
    finally:
        srv.close()
    gls.glsCvTerminate()
    gls.glsCvInit(0)


def test_release_does_not_wait_for_other_streams(gls):
    hip = _hip()
    s_long, s_short = _nonblocking_stream(hip), _nonblocking_stream(hip)
    try:
        big = gls.GlsMat(_rand_c64(4096, 2))
        # about 10 ms of in-place transforms on s_long (handle calls: the
        # library sees every use of `big`)
        for i in range(100):
            gls.fft(big, big, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE if i & 1 else 0, s_long)
        small = gls.GlsMat(_rand_c64(256, 3))
        gls.fft(small, small, 0, s_short)
        assert hip.hipStreamSynchronize(ctypes.c_void_p(s_short)) == 0
        small.release()  # last reference, used on s_short only
        # the release returned while s_long still runs: it did not wait for the device
        assert hip.hipStreamQuery(ctypes.c_void_p(s_long)) == HIP_ERROR_NOT_READY
        big.release(s_long)  # stream-ordered: never waits
        assert hip.hipStreamQuery(ctypes.c_void_p(s_long)) == HIP_ERROR_NOT_READY
    finally:
        _destroy(hip, s_long)
        _destroy(hip, s_short)


def test_released_block_not_reused_before_its_stream_finishes(gls):
    hip = _hip()
    s1, s2 = _nonblocking_stream(hip), _nonblocking_stream(hip)
    try:
        n = 1024
        busy = gls.GlsMat(_rand_c64(4096, 5))
        x = gls.GlsMat(_rand_c64(n, 4))
        # s1: about 20 ms of transforms of another buffer, then x's own, so x's
        # uses are still queued while the next buffers are made and filled
        for i in range(200):
            gls.fft(busy, busy, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE if i & 1 else 0, s1)
        for i in range(10):
            gls.fft(x, x, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE if i & 1 else 0, s1)
        x.release()
        # same size: a reuse of x's block now would be overwritten by s1's
        # transforms after the upload below (s1 and s2 do not order each other)
        y = gls.GlsMat((n, n), gls.CV_32FC2)
        pattern = np.arange(n * n * 2, dtype=np.float32).reshape(n, n, 2)
        y.upload(pattern, stream=s2)
        assert hip.hipStreamQuery(ctypes.c_void_p(s1)) == HIP_ERROR_NOT_READY
        assert hip.hipStreamSynchronize(ctypes.c_void_p(s1)) == 0
        assert np.array_equal(y.download(stream=s2), pattern)
        # a block released with its stream is reused once that stream passed it
        # (a size no other buffer here has, so the cache holds one such block)
        u = gls.GlsMat((1000, 1234), gls.CV_32FC2)
        u_ptr = u.dev_ptr
        u.release(s2)
        assert hip.hipStreamSynchronize(ctypes.c_void_p(s2)) == 0
        w = gls.GlsMat((1000, 1234), gls.CV_32FC2)
        assert w.dev_ptr == u_ptr
        tail = np.ones((1000, 1234, 2), np.float32)
        w.upload(tail)
        assert np.array_equal(w.download(), tail)
    finally:
        _destroy(hip, s1)
        _destroy(hip, s2)
