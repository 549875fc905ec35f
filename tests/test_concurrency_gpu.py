"""Several host threads, each on its own HIP stream, calling the C ABI at the
same time: every result equals, bit for bit, the result of the same call made
alone on the default stream.

The reference binds one GL context to the thread that called glsCvInit and
keeps unguarded shader singletons (glsCV/glsCV.cpp:130-136, glsFft.cpp:83,197),
so it has no concurrent callers; this library states that calls on distinct
streams are thread-safe (shared per-device state is initialised once under a
lock, per-stream scratch buffers are created under a lock).  ctypes releases
the GIL during each foreign call, so the threads run in the library together.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(256, 256, 1), (512, 256, 2), (1024, 1024, 3), (128, 2048, 4)]
REPS = 8


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.fail("libamdhip64.so not found")


def _inputs(rows, cols, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).astype(np.complex64)
    a = rng.uniform(0, 1, (rows, cols)).astype(np.float32)
    b = np.roll(a, (rows // 8 - seed, seed - cols // 8), axis=(0, 1)).astype(np.float32)
    return x, a, b


def _run(gls, x, a, b, stream):
    """FFT (SHIFT), the fused spectrum view and a phase correlation on `stream`."""
    src = gls.GlsMat(x)
    spec = gls.fft(src, None, gls.GLS_FFT_SHIFT, stream)
    y = spec.download(stream)
    view = gls.spectrumView(src, None, gls.GLS_FFT_SHIFT, 1.0, 0.0, 1.0, stream).download(stream)
    poc = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b), None, stream)
    return y, view, poc


def test_threads_on_own_streams_match_serial(gls):
    hip = _hip()
    inputs = [_inputs(*c) for c in CASES]
    serial = [_run(gls, *inp, None) for inp in inputs]
    for y, _, (_, _, peak) in serial:
        assert np.isfinite(y).all()
        assert peak is not None

    errors = []
    start = threading.Barrier(len(CASES))

    def worker(k):
        s = ctypes.c_void_p()
        try:
            if hip.hipStreamCreate(ctypes.byref(s)) != 0:
                errors.append(f"thread {k}: hipStreamCreate failed")
                return
            start.wait(timeout=60)
            for r in range(REPS):
                y, view, poc = _run(gls, *inputs[k], s.value)
                y0, view0, poc0 = serial[k]
                if not np.array_equal(y.view(np.uint32), y0.view(np.uint32)):
                    errors.append(f"thread {k} rep {r}: FFT bits differ from the serial call")
                if not np.array_equal(view.view(np.uint32), view0.view(np.uint32)):
                    errors.append(f"thread {k} rep {r}: spectrum view bits differ from the serial call")
                if poc != poc0:
                    errors.append(f"thread {k} rep {r}: phase correlation {poc} != serial {poc0}")
        except Exception as e:  # reported by the main thread
            errors.append(f"thread {k}: {type(e).__name__}: {e}")
        finally:
            if s.value:
                hip.hipStreamSynchronize(s)
                hip.hipStreamDestroy(s)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(CASES))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a worker thread did not finish"
    assert not errors, "\n".join(errors[:10])


def test_phase_correlations_on_four_streams_overlap(gls):
    """Four threads on their own streams finish a fixed set of phase
    correlations in less wall time than the same calls made one after another
    on one stream: each stream has its own workspace, so no lock is held
    across a call's synchronisation and the chains overlap on the device."""
    hip = _hip()
    n, reps, k = 256, 60, 4
    pairs = []
    for t in range(k):
        _, a, b = _inputs(n, n, 11 + t)
        pairs.append((gls.GlsMat(a), gls.GlsMat(b)))
    ptrs = [(a.dev_ptr, b.dev_ptr) for a, b in pairs]
    streams = []
    for _ in range(k):
        s = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
        streams.append(s.value)
    try:
        ref = [gls.phase_correlate_dev(pa, pb, None, n, n, streams[t]) for t, (pa, pb) in enumerate(ptrs)]
        for t, (pa, pb) in enumerate(ptrs):  # warm every stream's workspace
            gls.phase_correlate_dev(pa, pb, None, n, n, streams[t])

        t0 = time.perf_counter()
        for _ in range(reps):
            for t, (pa, pb) in enumerate(ptrs):
                r = gls.phase_correlate_dev(pa, pb, None, n, n, streams[0])
                assert (r.peak_x, r.peak_y) == (ref[t].peak_x, ref[t].peak_y)
        serial = time.perf_counter() - t0

        errors = []
        start = threading.Barrier(k + 1)

        def worker(t):
            try:
                start.wait(timeout=60)
                pa, pb = ptrs[t]
                for _ in range(reps):
                    r = gls.phase_correlate_dev(pa, pb, None, n, n, streams[t])
                    if (r.peak_x, r.peak_y, r.shift_x, r.shift_y, r.response) != (
                            ref[t].peak_x, ref[t].peak_y, ref[t].shift_x, ref[t].shift_y, ref[t].response):
                        errors.append(f"thread {t}: result differs from the first call")
            except Exception as e:  # reported by the main thread
                errors.append(f"thread {t}: {type(e).__name__}: {e}")

        threads = [threading.Thread(target=worker, args=(t,)) for t in range(k)]
        for th in threads:
            th.start()
        start.wait(timeout=60)
        t0 = time.perf_counter()
        for th in threads:
            th.join(timeout=100)
        concurrent = time.perf_counter() - t0
        assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
        assert not errors, "\n".join(errors[:10])
        print(f"{k * reps} phase correlations {n}x{n}: serial {serial * 1e3:.1f} ms, "
              f"{k} streams {concurrent * 1e3:.1f} ms")
        assert concurrent < 0.9 * serial, (serial, concurrent)
    finally:
        for s in streams:
            hip.hipStreamSynchronize(ctypes.c_void_p(s))
            hip.hipStreamDestroy(ctypes.c_void_p(s))
