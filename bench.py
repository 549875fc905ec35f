#!/usr/bin/env python3
"""Benchmark of the MI355X 2D FFT / phase-correlation path.

Workload (one "step"): forward + inverse(SCALE) 2D FFT of a 4096x4096
interleaved complex-float array (entries i.i.d. N(0,1), seeded), resident in
HBM before the timed region.  value = algorithmic GB/s = 2 * 32*M*N bytes per
step / step time (32*M*N = two sweeps, each reading and writing M*N*8 bytes).

Also reported (one JSON line):
  roofline     the dominant pass kernel: algorithmic bytes per launch
               (16*M*N) / its average duration, timed with HIP events on the
               stream it runs on around back-to-back launches of that pass
               (per-launch event timing, which adds the launch gap, is
               reported beside it); traffic from rocprofv3 --pmc summaries
               when profiles/traffic.json holds them
  cpu_baseline the restated reference CPU Stockham (oracle) on a bounded
               sample of the same workload on this host's cores
  accuracy     rel-RMS of one forward transform against numpy's FP64 FFT
  poc          per-request phase correlation (plain path) at mixed sizes:
               req/s and p50/p99 latency
  sizes        one forward SCALE FFT per size 64..4096 and one phase
               correlation per size 64..2048 (the reference's benchmark sizes)
  hbm_copy     a device-to-device copy of the same buffer, as a measured
               bandwidth line beside the FFT (and the FFT step's ratio to it)
  out_of_cache the same step rotating over 4 src/dst pairs (1 GiB, four times
               the 256 MB last-level cache), each pass and a copy alike
               (value_out_of_cache, roofline.frac_out_of_cache); fft_8192 the
               step at 8192^2 (1 GiB of buffers)
  poc_serving  plain (new buffers per request), plain_reuse (preallocated,
               pinned uploads), plain_ops (the reference's chain one operation
               per launch), the server (burst, Poisson), the device-resident
               batching gain over both per-request chains, and the restated
               reference chain on the CPU (1 core per size, all cores on the
               mixed pool)
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8.0 TB/s
WARM_S = 0.2           # minimum untimed warmup (seconds of steps) before timing


class Hip:
    """The few HIP runtime calls the benchmark needs (events, streams)."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = ctypes.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")

    def ck(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with hipError {rc}")

    def stream(self):
        s = ctypes.c_void_p()
        self.ck(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def event(self):
        e = ctypes.c_void_p()
        self.ck(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def record(self, ev, stream):
        self.ck(self.L.hipEventRecord(ctypes.c_void_p(ev), ctypes.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        ms = ctypes.c_float()
        self.ck(self.L.hipEventSynchronize(ctypes.c_void_p(e1)), "hipEventSynchronize")
        self.ck(self.L.hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(e0), ctypes.c_void_p(e1)),
                "hipEventElapsedTime")
        return ms.value

    def sync(self):
        self.ck(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def capture(self, stream, fn):
        """Record what fn() launches on `stream` into an executable HIP graph
        (thread-local capture mode)."""
        self.ck(self.L.hipStreamBeginCapture(ctypes.c_void_p(stream), 1), "hipStreamBeginCapture")
        try:
            fn()
        finally:
            g = ctypes.c_void_p()
            rc = self.L.hipStreamEndCapture(ctypes.c_void_p(stream), ctypes.byref(g))
        self.ck(rc, "hipStreamEndCapture")
        ex = ctypes.c_void_p()
        self.ck(self.L.hipGraphInstantiate(ctypes.byref(ex), g, None, None, ctypes.c_size_t(0)),
                "hipGraphInstantiate")
        return ex.value, g.value

    def graph_launch(self, ex, stream):
        self.ck(self.L.hipGraphLaunch(ctypes.c_void_p(ex), ctypes.c_void_p(stream)), "hipGraphLaunch")

    def graph_destroy(self, ex, g):
        self.ck(self.L.hipGraphExecDestroy(ctypes.c_void_p(ex)), "hipGraphExecDestroy")
        self.ck(self.L.hipGraphDestroy(ctypes.c_void_p(g)), "hipGraphDestroy")


def bench_copy(gls, hip, stream, n, reps=5):
    """Copy-kernel calibration for the PMC traffic counters: device-to-device
    copies of one n x n c32 buffer (8*n*n bytes read and written each)."""
    a = gls.GlsMat(np.zeros((n, n), np.complex64))
    b = gls.GlsMat((n, n), gls.CV_32FC2)
    for _ in range(reps):
        gls.lib().glsx_buf_copy(a.handle, b.handle, stream)
    hip.sync()
    return {"mode": "copy", "bytes_per_copy": 8 * n * n, "copies": reps}


def copy_bandwidth(gls, hip, stream, n, reps=20):
    """Sanity line beside the FFT: device-to-device copies of one n x n c32
    buffer (glsx_buf_copy, back to back after a warmup); GB/s counts the bytes
    read and written, 2 * 8*n*n per copy."""
    a = gls.GlsMat(np.zeros((n, n), np.complex64))
    b = gls.GlsMat((n, n), gls.CV_32FC2)
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < WARM_S:
        for _ in range(5):
            gls.lib().glsx_buf_copy(a.handle, b.handle, stream)
        hip.sync()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(reps):
        gls.lib().glsx_buf_copy(a.handle, b.handle, stream)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1) / reps
    gbs = 2 * 8.0 * n * n / (ms * 1e-3) / 1e9
    return {"bytes_per_copy": 2 * 8 * n * n, "ms": round(ms, 5), "gbs": round(gbs, 1),
            "pct_hbm_peak": round(100.0 * gbs / HBM_PEAK_GBS, 2)}


def bench_fft(gls, hip, stream, n, steps, warmup, accuracy=True):
    rng = np.random.default_rng(20261203)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    a = gls.GlsMat(x)
    b = gls.GlsMat((n, n), gls.CV_32FC2)
    pa, pb = a.dev_ptr, b.dev_ptr
    C2 = gls.CV_32FC2

    def step():
        gls.fft_dev(pa, C2, pb, n, n, 0, stream)
        gls.fft_dev(pb, C2, pb, n, n, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE, stream)

    # untimed warmup: W steps, and at least WARM_S seconds of them -- the GPU's
    # clocks take some tens of milliseconds of load to reach their steady state
    t_w = time.perf_counter()
    done = 0
    while done < warmup or time.perf_counter() - t_w < WARM_S:
        for _ in range(10):
            step()
        done += 10
        hip.sync()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    hip.record(e1, stream)
    ms_total = hip.elapsed_ms(e0, e1)
    hip.sync()
    wall = time.perf_counter() - t0
    ms_step = ms_total / steps
    bytes_step = 2 * 32.0 * n * n
    gbs = bytes_step / (ms_step * 1e-3) / 1e9
    flops_step = 2 * 5.0 * n * n * np.log2(n * n)

    # the same steps replayed from a captured HIP graph (10 steps per graph):
    # no host launch cost between kernels
    per_graph = 10
    ex, g = hip.capture(stream, lambda: [step() for _ in range(per_graph)])
    for _ in range(3):
        hip.graph_launch(ex, stream)
    hip.sync()
    n_graph = max(1, steps // per_graph)
    hip.record(e0, stream)
    for _ in range(n_graph):
        hip.graph_launch(ex, stream)
    hip.record(e1, stream)
    ms_step_graph = hip.elapsed_ms(e0, e1) / (n_graph * per_graph)
    hip.graph_destroy(ex, g)

    # per-pass launch durations, from HIP events on the stream the passes run
    # on.  The passes alternate forward and inverse(SCALE) so the data keeps
    # its magnitude (kernel time depends slightly on the data: overflowed or
    # zero buffers run faster than random ones).
    #   back to back: events around `reps` launches of one pass queued one
    #     after another, divided by `reps` -- the kernel's average duration,
    #     as a kernel trace measures it (the queue hides the launch gaps);
    #   single: events around every launch, averaged -- each value also
    #     holds the gap between the event and the kernel (a few microseconds).
    inv_sc = gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE
    row_pass = lambda fl: gls.fft1d_dev(pb, C2, pb, n, n, 1, n, fl, stream)
    col_pass = lambda fl: gls.fft1d_dev(pb, C2, pb, n, n, n, 1, fl, stream)

    def time_back_to_back(run, reps):
        e_a, e_b = hip.event(), hip.event()
        hip.record(e_a, stream)
        for i in range(reps):
            run(inv_sc if i & 1 else 0)
        hip.record(e_b, stream)
        return hip.elapsed_ms(e_a, e_b) / reps

    def time_single(run, reps):
        ev = [(hip.event(), hip.event()) for _ in range(reps)]
        for i in range(reps):
            hip.record(ev[i][0], stream)
            run(inv_sc if i & 1 else 0)
            hip.record(ev[i][1], stream)
        hip.sync()
        return float(np.mean([hip.elapsed_ms(*ev[i]) for i in range(reps)]))

    gls.lib().glsx_buf_copy(a.handle, b.handle, stream)
    reps = max(20, steps) & ~1
    ms_row, ms_col = time_back_to_back(row_pass, reps), time_back_to_back(col_pass, reps)
    ms_row1, ms_col1 = time_single(row_pass, reps), time_single(col_pass, reps)

    # accuracy of one forward transform against numpy's FP64 FFT
    relrms = None
    if accuracy:
        gls.fft_dev(pa, C2, pb, n, n, 0, stream)
        hip.sync()
        y = b.download_complex().astype(np.complex128)
        ref = np.fft.fft2(x.astype(np.complex128))
        relrms = float(np.linalg.norm((y - ref).ravel()) / np.linalg.norm(ref.ravel()))
    return dict(ms_step=ms_step, gbs=gbs, wall=wall, gflops=flops_step / (ms_step * 1e-3) / 1e9,
                ms_row=ms_row, ms_col=ms_col, ms_row1=ms_row1, ms_col1=ms_col1, relrms=relrms, ms_step_graph=ms_step_graph,
                gbs_graph=bytes_step / (ms_step_graph * 1e-3) / 1e9)


def bench_out_of_cache(gls, hip, stream, n, pairs=4, steps=40):
    """The headline step on a working set the 256 MB last-level cache cannot
    hold: `pairs` distinct (src, dst) pairs of n x n c32 buffers (at 4096^2,
    4 pairs = 1 GiB), step i running forward(src_k -> dst_k) + inverse(SCALE)
    in place on pair k = i mod pairs, so every step starts from data last
    touched `pairs - 1` steps (768 MiB) earlier.  Also each pass alone on the
    same rotation (in place on dst_k, alternating forward and inverse(SCALE)),
    and a device copy src_k -> dst_k on it, all with HIP events on `stream`."""
    rng = np.random.default_rng(20261203)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    srcs = [gls.GlsMat(x)]
    for _ in range(pairs - 1):
        srcs.append(srcs[0].clone())
    dsts = [gls.GlsMat((n, n), gls.CV_32FC2) for _ in range(pairs)]
    pa, pb = [m.dev_ptr for m in srcs], [m.dev_ptr for m in dsts]
    C2, INV = gls.CV_32FC2, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE

    def timed(fn, reps):
        t_w = time.perf_counter()
        i = 0
        while i < 2 * pairs or time.perf_counter() - t_w < WARM_S:
            fn(i)
            i += 1
            if i % pairs == 0:
                hip.sync()
        hip.sync()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for i in range(reps):
            fn(i)
        hip.record(e1, stream)
        return hip.elapsed_ms(e0, e1) / reps

    def step(i):
        k = i % pairs
        gls.fft_dev(pa[k], C2, pb[k], n, n, 0, stream)
        gls.fft_dev(pb[k], C2, pb[k], n, n, INV, stream)

    ms_step = timed(step, steps)
    ms_row = timed(lambda i: gls.fft1d_dev(pb[i % pairs], C2, pb[i % pairs], n, n, 1, n,
                                           INV if (i // pairs) & 1 else 0, stream), steps)
    ms_col = timed(lambda i: gls.fft1d_dev(pb[i % pairs], C2, pb[i % pairs], n, n, n, 1,
                                           INV if (i // pairs) & 1 else 0, stream), steps)
    ms_copy = timed(lambda i: gls.lib().glsx_buf_copy(srcs[i % pairs].handle, dsts[i % pairs].handle, stream), steps)
    gbs = 2 * 32.0 * n * n / (ms_step * 1e-3) / 1e9
    g = lambda ms: 16.0 * n * n / (ms * 1e-3) / 1e9
    return {"workload": f"forward + inverse(SCALE) {n}x{n} c32 rotating over {pairs} src/dst pairs "
                        f"({pairs * 16 * n * n >> 20} MiB touched)",
            "ms_per_step": round(ms_step, 5), "gbs": round(gbs, 1), "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
            "row_pass": {"ms": round(ms_row, 5), "gbs": round(g(ms_row), 1)},
            "col_pass": {"ms": round(ms_col, 5), "gbs": round(g(ms_col), 1),
                         "frac": round(g(ms_col) / HBM_PEAK_GBS, 4)},
            "device_copy": {"ms": round(ms_copy, 5), "gbs": round(g(ms_copy), 1)},
            "timing": "HIP events around back-to-back launches, each on the next pair"}


def bench_fft_big(gls, hip, stream, n=8192, reps=10):
    """forward + inverse(SCALE) of an n x n c32 array (8192^2: two 512 MiB
    buffers, twice the last-level cache), and each pass alone."""
    rng = np.random.default_rng(20261203)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    a, b = gls.GlsMat(x), gls.GlsMat((n, n), gls.CV_32FC2)
    pa, pb = a.dev_ptr, b.dev_ptr
    C2, INV = gls.CV_32FC2, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE
    del x

    def timed(fn):
        for i in range(3):
            fn(i)
        hip.sync()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for i in range(reps):
            fn(i)
        hip.record(e1, stream)
        return hip.elapsed_ms(e0, e1) / reps

    ms_step = timed(lambda i: (gls.fft_dev(pa, C2, pb, n, n, 0, stream), gls.fft_dev(pb, C2, pb, n, n, INV, stream)))
    # both transforms out of place (a -> b, b -> a); both forms run the
    # decimated column passes at 8192 rows (the column halves go to scratch)
    ms_oop = timed(lambda i: (gls.fft_dev(pa, C2, pb, n, n, 0, stream), gls.fft_dev(pb, C2, pa, n, n, INV, stream)))
    ms_row = timed(lambda i: gls.fft1d_dev(pb, C2, pb, n, n, 1, n, INV if i & 1 else 0, stream))
    ms_col = timed(lambda i: gls.fft1d_dev(pb, C2, pb, n, n, n, 1, INV if i & 1 else 0, stream))
    gbs = 2 * 32.0 * n * n / (ms_step * 1e-3) / 1e9
    gbs_oop = 2 * 32.0 * n * n / (ms_oop * 1e-3) / 1e9
    g = lambda ms: 16.0 * n * n / (ms * 1e-3) / 1e9
    return {"workload": f"forward + inverse(SCALE) {n}x{n} c32 (forward a -> b, inverse b -> b in place)",
            "ms_per_step": round(ms_step, 5),
            "gbs": round(gbs, 1), "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
            "out_of_place": {"workload": "forward a -> b, inverse(SCALE) b -> a", "ms_per_step": round(ms_oop, 5),
                             "gbs": round(gbs_oop, 1), "pct_hbm_peak": round(100 * gbs_oop / HBM_PEAK_GBS, 2)},
            "row_pass": {"ms": round(ms_row, 5), "gbs": round(g(ms_row), 1)},
            "col_pass": {"ms": round(ms_col, 5), "gbs": round(g(ms_col), 1),
                         "note": "one 8192-point column pass in place (the row-column path's kernel)"}}


def bench_passes(gls, hip, stream, n, reps):
    """The column pass and the row pass of an n x n complex array, each `reps`
    times back to back in place (1D transforms over the columns, then over the
    rows): for per-pass hardware counters."""
    x = (np.random.default_rng(1).standard_normal((n, n)) +
         1j * np.random.default_rng(2).standard_normal((n, n))).astype(np.complex64)
    a = gls.GlsMat(x)
    C2 = gls.CV_32FC2
    out = {}
    for name, es, bs in (("col_pass", n, 1), ("row_pass", 1, n)):
        e0, e1 = hip.event(), hip.event()
        gls.fft1d_dev(a.dev_ptr, C2, a.dev_ptr, n, n, es, bs, 0, stream)
        hip.record(e0, stream)
        for _ in range(reps):
            gls.fft1d_dev(a.dev_ptr, C2, a.dev_ptr, n, n, es, bs, 0, stream)
        hip.record(e1, stream)
        hip.sync()
        out[name] = {"ms": round(hip.elapsed_ms(e0, e1) / reps, 5)}
    return {"workload": f"{n}x{n} c32, each pass {reps} times in place", **out}


def bench_real_fft(gls, hip, stream, sizes=(1024, 4096), reps=50):
    """gls::fft of a real n x n frame (CV_32FC1 -> CV_32FC2, forward,
    unscaled), the glsCam input.  From 4096^2 up: two passes through
    half-size spectra, 20 algorithmic bytes per pixel; below, the complex
    transform of the frame, 28 (4 + 8 in the row pass, 16 in the column
    pass)."""
    out = {"workload": "forward 2D FFT of a real n x n float32 frame (uniform [0,1)) into the full complex spectrum"}
    for n in sizes:
        x = np.random.default_rng(6).uniform(0, 1, (n, n)).astype(np.float32)
        a, b = gls.GlsMat(x), gls.GlsMat((n, n), gls.CV_32FC2)
        pa, pb = a.dev_ptr, b.dev_ptr
        for _ in range(5):
            gls.fft_dev(pa, gls.CV_32FC1, pb, n, n, 0, stream)
        hip.sync()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for _ in range(reps):
            gls.fft_dev(pa, gls.CV_32FC1, pb, n, n, 0, stream)
        hip.record(e1, stream)
        ms = hip.elapsed_ms(e0, e1) / reps
        y = b.download_complex().astype(np.complex128) if n <= 1024 else None
        err = None
        if y is not None:
            ref = np.fft.fft2(x.astype(np.float64))
            err = float(np.linalg.norm((y - ref).ravel()) / np.linalg.norm(ref.ravel()))
        bpp = 20 if n >= 4096 else 28
        out[f"{n}x{n}"] = {"ms": round(ms, 5), "path": "half-size spectra" if bpp == 20 else "complex transform",
                           "bytes_per_px": bpp, "gbs": round(bpp * n * n / (ms * 1e-3) / 1e9, 1),
                           "relrms_vs_fp64": err}
    return out


def bench_split(gls, hip, stream, n=2048, reps=20, seed=4):
    """Split-precision MFMA variant beside the FP32 kernels: forward 2D FFT of
    n x n c32 holding 8 random-frequency tones (amplitudes log-uniform over
    1e-3..1e3) plus N(0, 1e-2) noise; launch time and error vs FP64 of each,
    over the whole spectrum and at the bins nearest the three weakest tones."""
    from glscv_amd.workloads import tones_2d

    x, tones = tones_2d(n, n, seed, return_tones=True)
    ref = np.fft.fft2(x.astype(np.complex128))
    weak = sorted(tones, key=lambda t: t[2])[:3]
    weak_bins = [(int(round(fy)) % n, int(round(fx)) % n, amp) for fy, fx, amp in weak]
    a = gls.GlsMat(x)
    b = gls.GlsMat((n, n), gls.CV_32FC2)
    pa, pb = a.dev_ptr, b.dev_ptr
    C2 = gls.CV_32FC2
    check_err, enabled = gls.split_mfma_check(n)
    res = {"workload": f"forward 2D FFT {n}x{n} c32, 8 tones (amplitudes 1e-3..1e3) + N(0,1e-2) noise",
           "fp64_check": {"len": n, "relrms": check_err, "enabled": enabled}}
    for name, flags in (("fp32", 0), ("split_mfma", gls.GLS_FFT_SPLIT_MFMA)):
        gls.fft_dev(pa, C2, pb, n, n, flags, stream)
        hip.sync()
        y = b.download_complex().astype(np.complex128)
        err = float(np.linalg.norm((y - ref).ravel()) / np.linalg.norm(ref.ravel()))
        weak_err = [{"bin": [r, c], "amplitude": float(f"{amp:.3g}"),
                     "relerr": float(f"{abs(y[r, c] - ref[r, c]) / abs(ref[r, c]):.3g}")} for r, c, amp in weak_bins]
        for _ in range(3):
            gls.fft_dev(pa, C2, pb, n, n, flags, stream)
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        for _ in range(reps):
            gls.fft_dev(pa, C2, pb, n, n, flags, stream)
        hip.record(e1, stream)
        ms = hip.elapsed_ms(e0, e1) / reps
        res[name] = {"ms": round(ms, 5), "gbs": round(32.0 * n * n / (ms * 1e-3) / 1e9, 1),
                     "relrms_vs_fp64": err, "weakest_tones_vs_fp64": weak_err}
    return res


def bench_spectrum_view(gls, hip, stream, sizes=(1024, 4096), reps=100):
    """glsCam's FFT view (glsCam.cpp:199-202) on an n x n real frame:
    fft(SHIFT) -> logMagSpectrums(1.0) -> normalize(0, 1, NORM_MINMAX) as the
    three separate calls (44 algorithmic bytes per pixel: the real-input FFT
    20, log magnitude 12, normalize 12), with the log magnitude fused into the
    FFT's column pass (28 B/px: row pass 8, column pass 8, normalize 12), and
    as one view call (three launches: the column pass also writes min/max
    partials, then the affine map; 24 B/px, and 20 B/px for frames of 4096^2
    and up, whose column pass stores half the plane and whose last pass
    writes both halves)."""
    res = {"workload": "n x n real frame: fft(SHIFT), logMagSpectrums(1.0), normalize(0,1,MINMAX)"}
    for n in sizes:
        rng = np.random.default_rng(5)
        x = rng.uniform(0, 1, (n, n)).astype(np.float32)
        frame = gls.GlsMat(x)
        spec = gls.GlsMat((n, n), gls.CV_32FC2)
        mag = gls.GlsMat((n, n), gls.CV_32FC1)
        pf, ps, pm = frame.dev_ptr, spec.dev_ptr, mag.dev_ptr
        SH, R32 = gls.GLS_FFT_SHIFT, gls.CV_32FC1

        def separate():
            gls.fft_dev(pf, R32, ps, n, n, SH, stream)
            gls.spectrum_dev(ps, pm, n * n, True, 1.0, stream)
            gls.normalize_dev(pm, pm, n * n, 0.0, 1.0, gls.NORM_MINMAX, stream)

        def fused():
            gls.fft_mag_dev(pf, R32, pm, n, n, SH, True, 1.0, stream)
            gls.normalize_dev(pm, pm, n * n, 0.0, 1.0, gls.NORM_MINMAX, stream)

        def view():
            gls.spectrum_view_dev(pf, R32, pm, n, n, SH, 1.0, 0.0, 1.0, stream)

        out = {}
        r = reps if n <= 1024 else max(10, reps // 5)
        view_bpp = 20 if n >= 4096 else 24
        for name, fn, bpp in (("separate", separate, 44), ("fused_log_mag", fused, 28), ("view", view, view_bpp)):
            for _ in range(5):
                fn()
            hip.sync()
            e0, e1 = hip.event(), hip.event()
            hip.record(e0, stream)
            for _ in range(r):
                fn()
            hip.record(e1, stream)
            ms = hip.elapsed_ms(e0, e1) / r
            out[name] = {"ms_per_frame": round(ms, 5), "frames_s": round(1e3 / ms, 1),
                         "gbs": round(bpp * n * n / (ms * 1e-3) / 1e9, 1), "bytes_per_px": bpp}
        out["view_speedup"] = round(out["separate"]["ms_per_frame"] / out["view"]["ms_per_frame"], 2)
        if n <= 1024:
            hip.sync()
            got = mag.download()
            spec64 = np.fft.fftshift(np.fft.fft2(x.astype(np.float64)))
            lm = np.log(np.abs(spec64) + 1.0)
            ref = (lm - lm.min()) / (lm.max() - lm.min())
            out["max_abs_err_vs_fp64"] = float(np.max(np.abs(got - ref)))
        res[f"{n}x{n}"] = out
    return res


def bench_batch(gls, hip, stream, count=256, seed=20261203, reps=5):
    """Batched uneven FFTs: `count` 2D FFTs, rows and cols each drawn from
    {64..2048}; one persistent hipGraph-captured launch (autotuned) against one
    launch pair per FFT on the same kernels."""
    rng = np.random.default_rng(seed)
    sizes = [64, 128, 256, 512, 1024, 2048]
    shapes = [(int(rng.choice(sizes)), int(rng.choice(sizes))) for _ in range(count)]
    srcs, dsts, descs = [], [], []
    for i, (r, c) in enumerate(shapes):
        x = (rng.uniform(-1, 1, (r, c)) + 1j * rng.uniform(-1, 1, (r, c))).astype(np.complex64)
        a, b = gls.GlsMat(x), gls.GlsMat((r, c), gls.CV_32FC2)
        srcs.append(a)
        dsts.append(b)
        descs.append(gls.FftDesc(a.dev_ptr, b.dev_ptr, r, c, gls.CV_32FC2, 0))
    t0 = time.perf_counter()
    plan = gls.Plan.fft_batch(descs, gls.PLAN_GRAPH | gls.PLAN_AUTOTUNE, keep=(srcs, dsts))
    plan_s = time.perf_counter() - t0
    info = plan.info()
    e0, e1 = hip.event(), hip.event()

    def timed(fn):
        fn()
        hip.sync()
        t = []
        for _ in range(reps):
            hip.record(e0, stream)
            fn()
            hip.record(e1, stream)
            t.append(hip.elapsed_ms(e0, e1))
        return float(np.median(t))

    ms_batch = timed(lambda: plan.execute(stream))
    plan.check()
    modes = {}
    for name, th, persistent in (("levels_256", 256, 0), ("levels_512", 512, 0), ("persistent_256", 256, 1),
                                 ("persistent_512", 512, 1), ("grouped_levels", 256, 2)):
        p = gls.Plan.fft_batch(descs, gls.PLAN_GRAPH, keep=(srcs, dsts), threads=th, persistent=persistent)
        modes[name] = round(timed(lambda: p.execute(stream)), 4)
        p.check()
        p.destroy()

    def one_by_one():
        for (r, c), a, b in zip(shapes, srcs, dsts):
            gls.fft_dev(a.dev_ptr, gls.CV_32FC2, b.dev_ptr, r, c, 0, stream)

    ms_plain = timed(one_by_one)
    total_bytes = sum(32.0 * r * c for r, c in shapes)
    # accuracy on a sample of the batch against numpy's FP64 FFT
    plan.execute(stream)
    hip.sync()
    errs = []
    for k in rng.choice(count, 12, replace=False):
        y = dsts[k].download_complex().astype(np.complex128)
        ref = np.fft.fft2(srcs[k].download_complex().astype(np.complex128))
        errs.append(float(np.linalg.norm(y - ref) / np.linalg.norm(ref)))
    return dict(count=count, shapes_from=sizes, ms_batch=round(ms_batch, 4), ms_one_launch_per_fft=round(ms_plain, 4),
                speedup=round(ms_plain / ms_batch, 2), batch_gbs=round(total_bytes / (ms_batch * 1e-3) / 1e9, 1),
                plain_gbs=round(total_bytes / (ms_plain * 1e-3) / 1e9, 1), max_relrms_sample=max(errs),
                launch={"threads": info.threads, "col_wcap": info.col_wcap, "persistent": info.persistent,
                        "levels": info.n_levels, "grid": info.grid, "items": info.n_items,
                        "graph": bool(info.uses_graph), "tuned": info.tuned},
                ms_by_launch_shape=modes,
                plan_build_s=round(plan_s, 3))


def _box5(img):
    acc = np.zeros_like(img, dtype=np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            acc += np.roll(img, (dy, dx), axis=(0, 1))
    return (acc / 25.0).astype(np.float32)


def _poc_pool(seed=31, per_size=3):
    """Request pool of the serving workload: uint8 noise / 255, 5x5 circular
    box blur, integer circular shift in [-N/4, N/4) plus N(0, 0.05) noise."""
    rng = np.random.default_rng(seed)
    pool = []
    for n in (256, 512, 1024, 2048):
        for _ in range(per_size):
            a = _box5(rng.integers(0, 256, (n, n), dtype=np.uint8).astype(np.float32) / 255.0)
            dx, dy = int(rng.integers(-n // 4, n // 4)), int(rng.integers(-n // 4, n // 4))
            b = (np.roll(a, (dy, dx), axis=(0, 1)) + rng.normal(0, 0.05, (n, n))).astype(np.float32)
            pool.append((a, b, n // 2 - dx, n // 2 - dy))
    return pool


def bench_serving(gls, hip, stream, n_req=400, seed=7):
    """Phase-correlation serving at mixed sizes {256,512,1024,2048}^2.

    plain:   one request at a time: upload both images, the chain with one
             launch per FFT pass, result download (synchronous).
    served:  PocServer (continuous batching, HBM-budget admission, spectra
             resident) under a saturating burst and under Poisson arrivals
             at 70% of the saturated rate; latency = submit -> result.
             served_pinned: the same with images already in pinned host
             memory (zero-copy submissions: no host copy into staging).
    device:  images already resident: per-request chain vs one batched plan
             over the same requests (the launch-level batching gain alone)."""
    rng = np.random.default_rng(seed)
    pool = _poc_pool()
    order = [int(k) for k in rng.integers(0, len(pool), n_req)]

    # --- plain, end to end ---------------------------------------------------
    n_plain = min(n_req, 120)
    lat, ok = [], True
    t0 = time.perf_counter()
    for k in order[:n_plain]:
        a, b, px, py = pool[k]
        t = time.perf_counter()
        ma, mb = gls.GlsMat(a), gls.GlsMat(b)
        r = gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, None, a.shape[0], a.shape[1], stream)
        lat.append(time.perf_counter() - t)
        ok &= (r.peak_x, r.peak_y) == (px, py)
    plain_s = time.perf_counter() - t0
    plain = dict(req_s=round(n_plain / plain_s, 1), p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 3),
                 p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 3), n_req=n_plain, peaks_exact=bool(ok))

    # the same pool in pinned host memory, for zero-copy submissions (images
    # that arrive in pinned buffers, e.g. from a capture device)
    pinned_pool = []
    for a, b, _, _ in pool:
        ha, hb = gls.HostBuffer(a.shape), gls.HostBuffer(b.shape)
        ha.array[...] = a
        hb.array[...] = b
        pinned_pool.append((ha, hb))

    # --- plain with reused buffers ------------------------------------------------
    # one request at a time as above, but with device buffers allocated once per
    # size and the images uploaded from pinned memory, as the server uploads
    # them: the per-request cost without allocation
    dev = {}
    for n in sorted({a.shape[0] for a, _, _, _ in pool}):
        ma, mb = gls.GlsMat((n, n), gls.CV_32FC1), gls.GlsMat((n, n), gls.CV_32FC1)
        dev[n] = (ma, mb, ma.dev_ptr, mb.dev_ptr)
    for k in range(len(pool)):  # warm every size's workspace
        n = pool[k][0].shape[0]
        gls.phase_correlate_dev(dev[n][2], dev[n][3], None, n, n, stream)
    lat, ok = [], True
    t0 = time.perf_counter()
    for k in order:
        a, b, px, py = pool[k]
        n = a.shape[0]
        ma, mb, pa, pb = dev[n]
        t = time.perf_counter()
        ma.upload_async(pinned_pool[k][0], stream)
        mb.upload_async(pinned_pool[k][1], stream)
        r = gls.phase_correlate_dev(pa, pb, None, n, n, stream)
        lat.append(time.perf_counter() - t)
        ok &= (r.peak_x, r.peak_y) == (px, py)
    reuse_s = time.perf_counter() - t0
    plain_reuse = dict(req_s=round(n_req / reuse_s, 1), p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 3),
                       p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 3), n_req=n_req, peaks_exact=bool(ok))

    # --- served ---------------------------------------------------------------

    def run_server(arrivals, pinned=False):
        # batches of at most 16 with 3 in flight: measured best for this mix (pinned
        # bursts 4,300 req/s against 3,700 with batches of 64)
        srv = gls.PocServer(hbm_budget_bytes=8 << 30, max_batch=16, max_inflight=3)
        ids = {}
        t_start = time.perf_counter()
        for k, t_arr in zip(order, arrivals):
            now = time.perf_counter() - t_start
            if t_arr > now:
                time.sleep(t_arr - now)
            if pinned:
                ids[srv.submit_pinned(*pinned_pool[k])] = k
            else:
                a, b, _, _ = pool[k]
                ids[srv.submit(a, b)] = k
        got = {}
        t_wait = time.perf_counter()
        while len(got) < len(ids):
            for c in srv.poll(256, timeout_ms=200):
                got[c.id] = c
            if time.perf_counter() - t_wait > 120:
                raise RuntimeError(f"serving: only {len(got)} of {len(ids)} requests completed")
        el = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        lat_ms = np.array([got[i].latency_s for i in ids]) * 1e3
        exact = all(got[i].status == 0 and (got[i].result.peak_x, got[i].result.peak_y) == pool[k][2:]
                    for i, k in ids.items())
        return dict(req_s=round(len(ids) / el, 1), p50_ms=round(float(np.percentile(lat_ms, 50)), 3),
                    p99_ms=round(float(np.percentile(lat_ms, 99)), 3), n_req=len(ids), peaks_exact=bool(exact),
                    batches=int(st.batches), mean_batch=round(st.mean_batch, 2),
                    peak_hbm_mib=round(st.peak_bytes / 2 ** 20, 1))

    burst = run_server([0.0] * n_req)
    lam = 0.7 * burst["req_s"]
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
    poisson = run_server(list(arrivals))
    poisson["rate_req_s"] = round(lam, 1)
    burst_p = run_server([0.0] * n_req, pinned=True)
    lam_p = 0.7 * burst_p["req_s"]
    poisson_p = run_server(list(np.cumsum(rng.exponential(1.0 / lam_p, n_req))), pinned=True)
    poisson_p["rate_req_s"] = round(lam_p, 1)

    # --- device-resident: per-request chain vs one batched plan ----------------
    mats = [(gls.GlsMat(a), gls.GlsMat(b)) for a, b, _, _ in pool]
    sub = order[:128]
    descs = [gls.PocDesc(mats[k][0].dev_ptr, mats[k][1].dev_ptr, None, pool[k][0].shape[0], pool[k][0].shape[1])
             for k in sub]
    ws_bytes = gls.Plan.poc_workspace_bytes(descs)
    ws = gls.GlsMat(((ws_bytes + (1 << 20) - 1) >> 20, 1 << 20), gls.CV_8UC1)
    plan = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, gls.PLAN_AUTOTUNE, keep=(mats, ws))
    plan.execute(stream)
    hip.sync()
    t_runs = []
    for _ in range(5):  # each run: the whole batch and its results on the host
        t0 = time.perf_counter()
        plan.execute(stream)
        res = plan.results(stream)
        t_runs.append(time.perf_counter() - t0)
    t_batch = float(np.median(t_runs))
    t0 = time.perf_counter()
    for k in sub:
        n = pool[k][0].shape[0]
        gls.phase_correlate_dev(mats[k][0].dev_ptr, mats[k][1].dev_ptr, None, n, n, stream)
    t_each = time.perf_counter() - t0
    exact = all((r.peak_x, r.peak_y) == pool[k][2:] for r, k in zip(res, sub))
    pi = plan.info()

    # --- the reference's chain, one operation per launch ---------------------------
    # glsPhaseCorrelate.cpp:190-209: fft(img1), fft(img2), mulSpectrumsPhaseOnly,
    # fft(INVERSE | SHIFT) in place, split (real plane), minMaxLoc, with buffers
    # allocated once per size.  "queued": every request's launches back to back,
    # each peak written to device memory (no host round trip: the launch-bound
    # floor of this path); "synced": each peak read on the host before the next
    # request, as the reference's minMaxLoc downloads it.  Integer peak only:
    # the reference's 5x5 centroid would add one more small download.
    ops = {}
    for n in sorted({a.shape[0] for a, _, _, _ in pool}):
        fa, fb, c = (gls.GlsMat((n, n), gls.CV_32FC2) for _ in range(3))
        sp = gls.GlsMat((n, n), gls.CV_32FC1)
        ops[n] = (fa, fb, c, sp)
    res_dev = gls.GlsMat((len(sub), 6), gls.CV_32SC1)          # one glsx_minmax_result per request

    def ops_chain(pa, pb, n, synced, slot=0):
        fa, fb, c, sp = ops[n]
        gls.fft_dev(pa, gls.CV_32FC1, fa.dev_ptr, n, n, 0, stream)
        gls.fft_dev(pb, gls.CV_32FC1, fb.dev_ptr, n, n, 0, stream)
        gls.mul_spectrums_dev(fa.dev_ptr, fb.dev_ptr, c.dev_ptr, n * n, gls.SPECTRUM_PHASE_ONLY, stream)
        gls.fft_dev(c.dev_ptr, gls.CV_32FC2, c.dev_ptr, n, n, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SHIFT, stream)
        gls.split_dev(c.dev_ptr, sp.dev_ptr, None, n * n, stream)
        if synced:
            return gls.minMaxLoc(sp, stream)[3]
        gls.min_max_loc_dev(sp.dev_ptr, gls.CV_32FC1, n, n, res_dev.dev_ptr + 24 * slot, stream)
        return None

    for n in ops:  # warm every size
        ma, mb = mats[[k for k in range(len(pool)) if pool[k][0].shape[0] == n][0]]
        ops_chain(ma.dev_ptr, mb.dev_ptr, n, True)
    t0 = time.perf_counter()
    for j, k in enumerate(sub):
        ops_chain(mats[k][0].dev_ptr, mats[k][1].dev_ptr, pool[k][0].shape[0], False, j)
    hip.sync()
    t_ops_q = time.perf_counter() - t0
    rq = res_dev.download()
    exact_q = all((int(rq[j, 4]), int(rq[j, 5])) == pool[k][2:] for j, k in enumerate(sub))
    t0 = time.perf_counter()
    exact_s = True
    for k in sub:
        exact_s &= ops_chain(mats[k][0].dev_ptr, mats[k][1].dev_ptr, pool[k][0].shape[0], True) == pool[k][2:]
    t_ops_s = time.perf_counter() - t0
    device = dict(n_req=len(sub), batched_req_s=round(len(sub) / t_batch, 1),
                  per_request_req_s=round(len(sub) / t_each, 1), speedup=round(t_each / t_batch, 2),
                  plain_ops_queued_req_s=round(len(sub) / t_ops_q, 1),
                  plain_ops_synced_req_s=round(len(sub) / t_ops_s, 1),
                  batched_vs_plain_ops=round(t_ops_q / t_batch, 2),
                  batched_vs_plain_ops_synced=round(t_ops_s / t_batch, 2),
                  peaks_exact=bool(exact and exact_q and exact_s),
                  launch={"threads": pi.threads, "col_wcap": pi.col_wcap, "persistent": pi.persistent,
                          "tuned": pi.tuned})

    # the same chain from host images (pinned uploads, synced peak), beside plain_reuse
    lat, ok = [], True
    t0 = time.perf_counter()
    for k in order:
        a, b, px, py = pool[k]
        n = a.shape[0]
        ma, mb, pa, pb = dev[n]
        t = time.perf_counter()
        ma.upload_async(pinned_pool[k][0], stream)
        mb.upload_async(pinned_pool[k][1], stream)
        ok &= ops_chain(pa, pb, n, True) == (px, py)
        lat.append(time.perf_counter() - t)
    ops_s = time.perf_counter() - t0
    plain_ops = dict(req_s=round(n_req / ops_s, 1), p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 3),
                     p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 3), n_req=n_req, peaks_exact=bool(ok))
    return dict(sizes=[256, 512, 1024, 2048], plain=plain, plain_reuse=plain_reuse, plain_ops=plain_ops,
                served_burst=burst, served_poisson=poisson, served_pinned_burst=burst_p,
                served_pinned_poisson=poisson_p, device_resident=device,
                served_vs_plain=round(burst["req_s"] / plain["req_s"], 2),
                served_pinned_vs_plain=round(burst_p["req_s"] / plain["req_s"], 2),
                served_vs_plain_reuse=round(burst["req_s"] / plain_reuse["req_s"], 2),
                served_pinned_vs_plain_reuse=round(burst_p["req_s"] / plain_reuse["req_s"], 2),
                served_pinned_vs_plain_ops=round(burst_p["req_s"] / plain_ops["req_s"], 2),
                batching_gain_device_resident=device["speedup"],
                batched_vs_plain_ops_device_resident=device["batched_vs_plain_ops"],
                note="plain: new GlsMats (allocation + pageable upload) per request, the reference's way; "
                     "plain_reuse: preallocated buffers, pinned uploads, the fused chain per request; "
                     "plain_ops: the reference's chain one operation per launch (fft, fft, "
                     "mulSpectrumsPhaseOnly, fft(INVERSE|SHIFT), split, minMaxLoc), preallocated buffers; "
                     "device_resident: inputs already in HBM -- speedup = batched plan vs the fused chain per "
                     "request, batched_vs_plain_ops = batched plan vs the queued one-operation-per-launch chain, "
                     "batched_vs_plain_ops_synced = the same with each peak read on the host as minMaxLoc does")


def bench_poc_plan(gls, hip, stream, n_req=400, seed=7, reps=20):
    """The device-resident batched plan of the serving workload alone (the
    first 128 requests of bench_serving's order, images in HBM, autotuned
    launch shape, captured graph): median of `reps` runs, each the whole
    batch with its results on the host (for profiling the plan's launches)."""
    rng = np.random.default_rng(seed)
    pool = _poc_pool()
    order = [int(k) for k in rng.integers(0, len(pool), n_req)]
    sub = order[:128]
    mats = [(gls.GlsMat(a), gls.GlsMat(b)) for a, b, _, _ in pool]
    descs = [gls.PocDesc(mats[k][0].dev_ptr, mats[k][1].dev_ptr, None, pool[k][0].shape[0], pool[k][0].shape[1])
             for k in sub]
    ws_bytes = gls.Plan.poc_workspace_bytes(descs)
    ws = gls.GlsMat(((ws_bytes + (1 << 20) - 1) >> 20, 1 << 20), gls.CV_8UC1)
    plan = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, gls.PLAN_GRAPH | gls.PLAN_AUTOTUNE, keep=(mats, ws))
    plan.execute(stream)
    hip.sync()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.execute(stream)
        res = plan.results(stream)
        t.append(time.perf_counter() - t0)
    exact = all((r.peak_x, r.peak_y) == pool[k][2:] for r, k in zip(res, sub))
    pi = plan.info()
    ms = float(np.median(t)) * 1e3
    nbytes = sum(60.0 * pool[k][0].size for k in sub)
    return dict(n_req=len(sub), ms_per_batch=round(ms, 3), req_s=round(len(sub) / ms * 1e3, 1),
                gbs=round(nbytes / (ms * 1e-3) / 1e9, 1), peaks_exact=bool(exact),
                launch={"threads": pi.threads, "col_wcap": pi.col_wcap, "persistent": pi.persistent,
                        "levels": pi.n_levels, "items": pi.n_items, "tuned": pi.tuned})


def bench_sizes(gls, hip, stream, reps=100, seed=12):
    """Per-size lines at the reference's benchmark sizes: one forward 2D FFT
    with SCALE on a resident N x N array (unittest_glsFft.cpp FFT_{64..2048}_
    SCALE_benchmark, plus 4096), HIP events over `reps` back-to-back calls;
    and one phase correlation of resident N x N images through the per-request
    chain, synchronous, wall clock (UnitTest_glsPhaseCorrelate.cpp benchmark
    sizes 64..2048)."""
    rng = np.random.default_rng(seed)
    fft_us, fft_gbs, poc_ms = {}, {}, {}
    e0, e1 = hip.event(), hip.event()
    for n in (64, 128, 256, 512, 1024, 2048, 4096):
        x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
        a, b = gls.GlsMat(x), gls.GlsMat((n, n), gls.CV_32FC2)
        C2 = gls.CV_32FC2
        # frames up to 256^2 are bound by the host's launch cost: a longer
        # warm-up and more calls steady the host side (the first size follows
        # the multithreaded serving leg)
        warm, r = (200, 4 * reps) if n <= 256 else (10, reps)
        for _ in range(warm):
            gls.fft_dev(a.dev_ptr, C2, b.dev_ptr, n, n, gls.GLS_FFT_SCALE, stream)
        hip.sync()
        hip.record(e0, stream)
        for _ in range(r):
            gls.fft_dev(a.dev_ptr, C2, b.dev_ptr, n, n, gls.GLS_FFT_SCALE, stream)
        hip.record(e1, stream)
        hip.sync()
        us = hip.elapsed_ms(e0, e1) / r * 1e3
        fft_us[n] = round(us, 2)
        fft_gbs[n] = round(32.0 * n * n / (us * 1e-6) / 1e9, 1)
    for n in (64, 128, 256, 512, 1024, 2048):
        img = rng.uniform(0, 1, (n, n)).astype(np.float32)
        ma, mb = gls.GlsMat(img), gls.GlsMat(np.roll(img, (3, -5), axis=(0, 1)))
        for _ in range(3):
            gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, None, n, n, stream)
        t0 = time.perf_counter()
        k = 20
        for _ in range(k):
            r = gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, None, n, n, stream)
        poc_ms[n] = round((time.perf_counter() - t0) / k * 1e3, 3)
        if (r.peak_x, r.peak_y) != (n // 2 + 5, n // 2 - 3):
            raise RuntimeError(f"phase correlation {n}x{n}: wrong peak ({r.peak_x}, {r.peak_y})")
    return {"fft_scale_us": fft_us, "fft_scale_gbs": fft_gbs, "poc_request_ms": poc_ms,
            "note": "fft: one forward SCALE transform per call, resident; poc: resident images, "
                    "per-request chain with result on the host"}


def cpu_baseline(n, budget_s=12.0, nthreads=1):
    """The restated reference CPU Stockham (oracle) on a bounded sample.  The
    reference runs it single-threaded (its OpenMP loops are commented out);
    nthreads > 1 splits rows / columns over that many host threads."""
    import oracle

    rng = np.random.default_rng(1)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and done < 64:
        oracle.fft_r4_f32(x, 0 if done % 2 == 0 else 3, nthreads=nthreads)
        done += 1
    el = time.perf_counter() - t0
    how = "single-threaded" if nthreads == 1 else f"{nthreads}-thread"
    total, allowed = host_cpus()
    return dict(value=round(done * 32.0 * n * n / el / 1e9, 4), unit="GB/s", cores=nthreads, kind="port",
                host_logical_cpus=total, host_cpus_allowed=allowed,
                sample=f"{done} {how} {n}x{n} c32 transforms (fwd/inv alternating) of the restated "
                       f"fft_stockhum.cpp radix-4 Stockham, {el:.1f} s")


def host_cpus():
    """(logical CPUs of the machine, CPUs this process may run on)."""
    total = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = total
    return total, allowed


def host_threads():
    """Host threads to use for the multi-threaded CPU lines: the CPUs this
    process may run on, at most 16 (the host share of one GPU)."""
    return max(1, min(16, host_cpus()[1]))


def cpu_poc_baseline(budget_single_s=4.0, budget_all_s=5.0):
    """The restated reference chain (oracle.phase_correlate: three restated
    radix-4 Stockham FFTs, POC, argmax, centroid; glsPhaseCorrelate.cpp:142-231)
    on the serving workload's images: single-threaded ms per request at each
    size, as the reference runs it, and requests in parallel on host_threads()
    threads (one request per thread at a time) for req/s on the mixed pool."""
    import concurrent.futures as cf

    import oracle

    pool = _poc_pool(per_size=1)
    ms, ok = {}, True
    per_size = budget_single_s / 4
    for a, b, px, py in pool:
        n = a.shape[0]
        t, done = [], 0
        t_start = time.perf_counter()
        while done < 1 or (time.perf_counter() - t_start < per_size and done < 20):
            t0 = time.perf_counter()
            r, _ = oracle.phase_correlate(a, b, None, nthreads=1)
            t.append(time.perf_counter() - t0)
            ok &= (r.peak_x, r.peak_y) == (px, py)
            done += 1
        ms[n] = round(float(np.median(t)) * 1e3, 3)
    total, allowed = host_cpus()
    th = host_threads()
    rng = np.random.default_rng(3)

    def one(k):
        a, b, px, py = pool[k]
        r, _ = oracle.phase_correlate(a, b, None, nthreads=1)
        return (r.peak_x, r.peak_y) == (px, py)

    done = 0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=th) as ex:
        while time.perf_counter() - t0 < budget_all_s:
            ks = [int(k) for k in rng.integers(0, len(pool), 2 * th)]
            for good in ex.map(one, ks):
                ok &= good
                done += 1
    el = time.perf_counter() - t0
    mean_ms = float(np.mean(list(ms.values())))
    return {"ms_per_request_1_core": ms, "req_s_1_core_mixed": round(1e3 / mean_ms, 2),
            "req_s_all_cores_mixed": round(done / el, 2), "cores": th, "host_logical_cpus": total,
            "host_cpus_allowed": allowed, "kind": "port", "peaks_exact": bool(ok),
            "sample": f"1-core: >=1 request per size (median); all cores: {done} requests of the mixed pool "
                      f"on {th} threads in {el:.1f} s; restated glsPhaseCorrelate.cpp chain"}


def cpu_numpy_fft(n, budget_s=3.0):
    """An extra line, not the reference's path: numpy.fft.fft2 / ifft2 (pocketfft,
    single-threaded) on the same n x n c32 input, GB/s (32*M*N per transform)."""
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    done, dtype = 0, None
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and done < 16:
        y = np.fft.fft2(x) if done % 2 == 0 else np.fft.ifft2(x)
        dtype = str(y.dtype)
        done += 1
    el = time.perf_counter() - t0
    return dict(value=round(done * 32.0 * n * n / el / 1e9, 4), unit="GB/s", cores=1,
                kind="numpy pocketfft (extra line, not the reference's path)",
                sample=f"{done} single-threaded {n}x{n} numpy.fft.fft2/ifft2 calls on complex64 input "
                       f"(computed in {dtype}), {el:.1f} s")


def cpu_fft_sizes(sizes=(256, 1024), budget_s=1.5):
    """Single-threaded restated Stockham at the small sizes, GB/s (32*M*N per transform)."""
    return {f"{n}x{n}": cpu_baseline(n, budget_s=budget_s) for n in sizes}


def load_traffic(n):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(f"{n}x{n}")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-poc", action="store_true", help="skip the phase-correlation serving leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven FFT leg")
    ap.add_argument("--no-acc", action="store_true", help="skip the accuracy check")
    ap.add_argument("--no-split", action="store_true",
                    help="skip the split-precision MFMA and spectrum-view legs")
    ap.add_argument("--no-sizes", action="store_true",
                    help="skip the per-size FFT and phase-correlation lines (64..4096)")
    ap.add_argument("--no-ooc", action="store_true",
                    help="skip the out-of-cache (rotating buffers) and 8192^2 lines")
    ap.add_argument("--mode", choices=["full", "copy", "real", "view", "batch", "passes"], default="full",
                    help="copy: only the copy-kernel calibration used to scale PMC traffic counters; "
                         "real: only the real-input FFT leg at --size; view: only the glsCam view leg at "
                         "--size; passes: the column pass and the row pass of an --size square, each "
                         "--steps times in place; batch: only the batched uneven FFT "
                         "leg and the device-resident phase-correlation plan (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import glscv_amd as gls

    gls.glsCvInit(0)
    hip = Hip()
    stream = hip.stream()
    n = args.size
    if args.mode == "copy":
        print(json.dumps(bench_copy(gls, hip, stream, n)))
        return
    if args.mode == "real":
        print(json.dumps(bench_real_fft(gls, hip, stream, sizes=(n,), reps=max(1, args.steps))))
        return
    if args.mode == "view":
        print(json.dumps(bench_spectrum_view(gls, hip, stream, sizes=(n,), reps=max(1, args.steps))))
        return
    if args.mode == "passes":
        print(json.dumps(bench_passes(gls, hip, stream, n, max(1, args.steps))))
        return
    if args.mode == "batch":
        print(json.dumps({"batched_uneven": bench_batch(gls, hip, stream),
                          "poc_plan_device_resident": bench_poc_plan(gls, hip, stream)}))
        return
    def leg(name):  # one progress line per leg on stderr; stdout keeps the one JSON line
        print(f"bench: {name}", file=sys.stderr, flush=True)

    leg(f"fft {n}x{n}")
    r = bench_fft(gls, hip, stream, n, args.steps, args.warmup, accuracy=not args.no_acc)
    copy = copy_bandwidth(gls, hip, stream, n)
    leg("out of cache")
    ooc = None if args.no_ooc else bench_out_of_cache(gls, hip, stream, n)
    leg("fft 8192x8192")
    big = None if args.no_ooc or n >= 8192 else bench_fft_big(gls, hip, stream, 8192)
    leg("split precision")
    split = None if args.no_split else bench_split(gls, hip, stream)
    leg("spectrum view")
    view = None if args.no_split else bench_spectrum_view(gls, hip, stream)
    leg("real-input fft")
    real = None if args.no_split else bench_real_fft(gls, hip, stream)
    leg("batched uneven ffts")
    batch = None if args.no_batch else bench_batch(gls, hip, stream)
    leg("phase-correlation serving")
    poc = None if args.no_poc else bench_serving(gls, hip, stream)
    leg("per-size lines")
    sizes = None if args.no_sizes else bench_sizes(gls, hip, stream)
    small = None
    if n != 1024:
        leg("fft 1024x1024")
        rs = bench_fft(gls, hip, stream, 1024, args.steps, args.warmup, accuracy=not args.no_acc)
        small = {"workload": "forward + inverse(SCALE) 2D complex FFT, 1024x1024 c32", "gbs": round(rs["gbs"], 1),
                 "pct_hbm_peak": round(100.0 * rs["gbs"] / HBM_PEAK_GBS, 2), "ms_per_step": round(rs["ms_step"], 5),
                 "graph_replay": {"gbs": round(rs["gbs_graph"], 1), "ms_per_step": round(rs["ms_step_graph"], 5),
                                  "pct_hbm_peak": round(100.0 * rs["gbs_graph"] / HBM_PEAK_GBS, 2)},
                 # each pass back to back on its own output (in place: every pass reads what the same
                 # workgroups, on the same XCDs, wrote); in the step each pass reads what the other
                 # pass wrote from all XCDs
                 "passes_alone_ms": {"row_pass": round(rs["ms_row"], 5), "col_pass": round(rs["ms_col"], 5),
                                     "sum_of_four": round(2 * (rs["ms_row"] + rs["ms_col"]), 5)},
                 "relrms_vs_fp64": rs["relrms"]}
    leg("cpu baselines")
    cpu = None if args.no_cpu else cpu_baseline(n)
    cpu_mt = None if args.no_cpu else cpu_baseline(n, budget_s=6.0, nthreads=host_threads())
    cpu_small = None if args.no_cpu else cpu_fft_sizes()
    cpu_np = None if args.no_cpu else cpu_numpy_fft(n)
    if poc is not None and not args.no_cpu:
        poc["cpu_baseline"] = cpu_poc_baseline()
        poc["served_vs_cpu_all_cores"] = round(poc["served_burst"]["req_s"] /
                                               poc["cpu_baseline"]["req_s_all_cores_mixed"], 1)

    dom_col = r["ms_col"] >= r["ms_row"]
    dom_ms = r["ms_col"] if dom_col else r["ms_row"]
    achieved = 16.0 * n * n / (dom_ms * 1e-3) / 1e9
    traffic = load_traffic(n)
    out = {
        "metric": "2D FFT GB/s (% HBM peak) at 4096^2 c32; phase-correlate req/s + p99 latency",
        "value": round(r["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: complex N(0,1), seed 20261203",
        "config": {"workload": f"forward + inverse(SCALE) 2D complex FFT, {n}x{n} c32 interleaved, "
                               f"resident in device memory (two buffers of {8 * n * n >> 20} MiB, which the "
                               f"256 MB last-level cache can hold; out_of_cache rotates over 1 GiB)",
                   "rows": n, "cols": n, "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * r["gbs"] / HBM_PEAK_GBS, 2),
        "gflops": round(r["gflops"], 1),
        "relrms_vs_fp64": r["relrms"],
        "graph_replay": {"gbs": round(r["gbs_graph"], 1), "ms_per_step": round(r["ms_step_graph"], 5),
                         "pct_hbm_peak": round(100.0 * r["gbs_graph"] / HBM_PEAK_GBS, 2)},
        "value_out_of_cache": ooc["gbs"] if ooc else None,
        "out_of_cache": ooc,
        "fft_8192": big,
        "roofline": {
            "bound": "hbm",
            # the column kernel the library picks for an n x n complex transform
            "kernel": (f"fft_col_pair_kernel<{n}, {2 if n == 8192 else 4}, false, false>"
                       if dom_col and n in (1024, 2048, 4096, 8192)
                       else f"fft_pass_kernel<{n}, true>" if dom_col
                       else f"fft_row_kernel<{n}>" if n >= 2048 else f"fft_pass_kernel<{n}, false>"),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # the same pass kernel timed on buffers rotated out of the last-level cache
            "frac_out_of_cache": ooc["col_pass"]["frac"] if ooc and dom_col else None,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "per_unit": f"16*M*N bytes per pass launch ({16 * n * n} B)",
            "avg_ms": {"row_pass": round(r["ms_row"], 5), "col_pass": round(r["ms_col"], 5),
                       "timing": "HIP events around back-to-back launches of one pass on its stream"},
            "avg_ms_single_launch": {"row_pass": round(r["ms_row1"], 5), "col_pass": round(r["ms_col1"], 5)},
        },
        "hbm_copy": dict(copy, fft_step_vs_copy=round(r["gbs"] / copy["gbs"], 3),
                         workload=f"device-to-device copy of the {n}x{n} c32 buffer"),
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_mt,
        "cpu_baseline_sizes": cpu_small,
        "cpu_baseline_numpy": cpu_np,
        "fft_1024": small,
        "split_mfma": split,
        "spectrum_view": view,
        "real_fft": real,
        "batched_uneven": batch,
        "poc_serving": poc,
        "sizes": sizes,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
