"""GPU tests of the continuous-batching phase-correlation server: every
result equals the same request run alone (bit for bit), integer peaks equal
the ground-truth shift, admission respects the HBM budget, and requests that
can never fit are rejected instead of blocking the queue."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def box5(img):
    acc = np.zeros_like(img, dtype=np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            acc += np.roll(img, (dy, dx), axis=(0, 1))
    return (acc / 25.0).astype(np.float32)


def pair(n, seed):
    rng = np.random.default_rng(seed)
    a = box5(rng.integers(0, 256, (n, n), dtype=np.uint8).astype(np.float32) / 255.0)
    dx, dy = int(rng.integers(-n // 4, n // 4)), int(rng.integers(-n // 4, n // 4))
    b = (np.roll(a, (dy, dx), axis=(0, 1)) + rng.normal(0, 0.05, (n, n))).astype(np.float32)
    return a, b, dx, dy


def collect(srv, n, timeout_s=60.0):
    out = {}
    t0 = time.time()
    while len(out) < n and time.time() - t0 < timeout_s:
        for c in srv.poll(256, timeout_ms=100):
            out[c.id] = c
    assert len(out) == n, f"only {len(out)} of {n} completions"
    return out


def alone(gls, a, b):
    ma, mb = gls.GlsMat(a), gls.GlsMat(b)
    return gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, None, a.shape[0], a.shape[1])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_results_same_alone_and_in_the_mix(gls, mode):
    """Every execution mode of the batches (per level, persistent, per level
    and kernel) returns the bits of the request run alone."""
    sizes = [256, 512, 1024, 2048, 256, 256, 512, 1024, 256, 512, 2048, 256]
    reqs = [pair(n, 900 + i) for i, n in enumerate(sizes)]
    srv = gls.PocServer(hbm_budget_bytes=2 << 30, max_batch=16, max_inflight=2, persistent=mode)
    ids = [srv.submit(a, b) for a, b, _, _ in reqs]
    done = collect(srv, len(ids))
    st = srv.stats()
    srv.close()
    assert st.completed == len(ids) and st.rejected == 0
    for rid, (a, b, dx, dy), n in zip(ids, reqs, sizes):
        c = done[rid]
        assert c.status == 0
        r = c.result
        assert (r.peak_x, r.peak_y) == (n // 2 - dx, n // 2 - dy)
        ref = alone(gls, a, b)
        assert (r.peak_x, r.peak_y, r.shift_x, r.shift_y, r.response) == \
               (ref.peak_x, ref.peak_y, ref.shift_x, ref.shift_y, ref.response)


def test_budget_admission_and_rejection(gls):
    # 40 MiB: one 1024^2 request needs ~20 MiB of device memory (two images,
    # the packed spectrum, the surface), a 2048^2 one ~80 MiB and can never be
    # admitted
    srv = gls.PocServer(hbm_budget_bytes=40 << 20, max_batch=8, max_inflight=2)
    reqs = [pair(1024, 1000 + i) for i in range(6)] + [pair(2048, 1100)] + [pair(256, 1200 + i) for i in range(4)]
    ids = [srv.submit(a, b) for a, b, _, _ in reqs]
    done = collect(srv, len(ids))
    st = srv.stats()
    srv.close()
    assert done[ids[6]].status == -6                       # GLSX_ERR_BUDGET
    for i, (a, b, dx, dy) in enumerate(reqs):
        if i == 6:
            continue
        n = a.shape[0]
        assert done[ids[i]].status == 0
        assert (done[ids[i]].result.peak_x, done[ids[i]].result.peak_y) == (n // 2 - dx, n // 2 - dy)
    assert st.rejected == 1 and st.completed == len(ids) - 1
    assert st.peak_bytes <= (40 << 20)
    # at most one 1024^2 request fits at a time, so they were served in
    # several launches while memory was released and reused
    assert st.batches >= 6


def test_window_and_streaming_arrivals(gls):
    n = 256
    wy = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)
    win = np.outer(wy, wy).astype(np.float32)
    srv = gls.PocServer(hbm_budget_bytes=1 << 30, max_batch=32, max_inflight=3)
    reqs = [pair(n, 1300 + i) for i in range(40)]
    ids = []
    for a, b, _, _ in reqs:                                # arrivals while earlier batches run
        ids.append(srv.submit(a, b, win))
        time.sleep(0.0005)
    done = collect(srv, len(ids))
    srv.close()
    for rid, (a, b, dx, dy) in zip(ids, reqs):
        ma, mb, mw = gls.GlsMat(a), gls.GlsMat(b), gls.GlsMat(win)
        ref = gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, mw.dev_ptr, n, n)
        r = done[rid].result
        assert (r.peak_x, r.peak_y, r.shift_x, r.shift_y) == (ref.peak_x, ref.peak_y, ref.shift_x, ref.shift_y)


def test_bad_submissions(gls):
    srv = gls.PocServer(hbm_budget_bytes=64 << 20, max_batch=4, max_inflight=1)
    with pytest.raises(gls.GlsError):
        srv.submit(np.zeros((48, 48), np.float32), np.zeros((48, 48), np.float32))
    with pytest.raises(gls.GlsError):
        srv.submit(np.zeros((8192, 2), np.float32), np.zeros((8192, 2), np.float32))
    srv.close()
    with pytest.raises(gls.GlsError):
        gls.PocServer(hbm_budget_bytes=1 << 10)


def test_submit_pinned_matches_submit(gls):
    sizes = [256, 1024, 512, 2048, 256]
    reqs = [pair(n, 1300 + i) for i, n in enumerate(sizes)]
    bufs = []
    for a, b, _, _ in reqs:
        ha, hb = gls.HostBuffer(a.shape), gls.HostBuffer(b.shape)
        ha.array[...] = a
        hb.array[...] = b
        bufs.append((ha, hb))
    srv = gls.PocServer(hbm_budget_bytes=1 << 30, max_batch=8, max_inflight=2)
    ids_p = [srv.submit_pinned(ha, hb) for ha, hb in bufs]
    ids_c = [srv.submit(a, b) for a, b, _, _ in reqs]
    done = collect(srv, len(ids_p) + len(ids_c))
    assert not srv._held                                   # every pinned buffer released at completion
    srv.close()
    for ip, ic, (a, b, dx, dy) in zip(ids_p, ids_c, reqs):
        p, c = done[ip].result, done[ic].result
        n = a.shape[0]
        assert done[ip].status == 0 and (p.peak_x, p.peak_y) == (n // 2 - dx, n // 2 - dy)
        assert (p.peak_x, p.peak_y, p.shift_x, p.shift_y, p.response) == \
               (c.peak_x, c.peak_y, c.shift_x, c.shift_y, c.response)


def test_submit_pinned_rejects_pageable_memory(gls):
    import ctypes
    a, b, _, _ = pair(64, 1400)
    srv = gls.PocServer(hbm_budget_bytes=64 << 20, max_batch=4, max_inflight=1)
    rid = ctypes.c_longlong()
    rc = gls.lib().glsx_server_submit_pinned(srv._h, a.ctypes.data, b.ctypes.data, None, 64, 64, ctypes.byref(rid))
    srv.close()
    assert rc == -1


def test_async_transfers_through_pinned_memory(gls):
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((512, 256)) + 1j * rng.standard_normal((512, 256))).astype(np.complex64)
    h_in = gls.HostBuffer((512, 256), np.complex64)
    h_out = gls.HostBuffer((512, 256), np.complex64)
    h_in.array[...] = x
    m = gls.GlsMat((512, 256), gls.CV_32FC2)
    m.upload_async(h_in)
    gls.fft(m, m, 0)
    m.download_async(h_out)
    gls.stream_sync()
    ref = gls.fft(gls.GlsMat(x), None, 0).download_complex()
    assert np.array_equal(h_out.array, ref)
    with pytest.raises(gls.GlsError):
        m.upload_async(gls.HostBuffer((4, 4), np.complex64))  # wrong size
