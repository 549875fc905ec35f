"""The allocation-free phase correlation (glsx_phase_correlate_dev_noalloc)
against glsx_phase_correlate_dev on the same inputs: two calls in a row, the
second smaller, each equal to the allocating entry point bit for bit (same
kernels, same workspace layout; only the workspace's address differs).  Sizes
beyond the static workspace are rejected before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(n, seed, dx, dy):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 1, (n, n)).astype(np.float32)
    b = (np.roll(a, (dy, dx), axis=(0, 1)) + rng.normal(0, 0.05, (n, n))).astype(np.float32)
    return a, b


def _hann(n):
    w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / (n - 1))
    return np.sqrt(np.outer(w, w)).astype(np.float32)


def test_noalloc_matches_allocating_entry_point(gls):
    gls.glsCvInit(0)
    out = gls.GlsMat((1, 32), gls.CV_8UC1)  # one glsx_poc_result (32 bytes) in device memory
    for n, (dx, dy), windowed in ((512, (37, -20), True), (256, (-9, 14), False)):
        print(f"no-alloc phase correlation {n}x{n} shift ({dx},{dy}) window={windowed}", flush=True)
        a, b = _pair(n, n + 5, dx, dy)
        ma, mb = gls.GlsMat(a), gls.GlsMat(b)
        mw = gls.GlsMat(_hann(n)) if windowed else None
        wp = mw.dev_ptr if mw is not None else None
        ref = gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, wp, n, n)
        gls.phase_correlate_dev_noalloc(ma.dev_ptr, mb.dev_ptr, wp, n, n, out.dev_ptr)
        got = gls.PocResult.from_buffer_copy(out.download().tobytes())
        assert (got.peak_x, got.peak_y) == (ref.peak_x, ref.peak_y) == (n // 2 - dx, n // 2 - dy)
        assert (got.shift_x, got.shift_y, got.response) == (ref.shift_x, ref.shift_y, ref.response)


def test_noalloc_rejects_sizes_beyond_the_static_workspace(gls):
    gls.glsCvInit(0)
    a = gls.GlsMat((4096, 4096), gls.CV_32FC1)
    out = gls.GlsMat((1, 32), gls.CV_8UC1)
    with pytest.raises(gls.GlsError) as e:
        gls.phase_correlate_dev_noalloc(a.dev_ptr, a.dev_ptr, None, 4096, 4096, out.dev_ptr)
    assert e.value.status == -2 and "static" in str(e.value)
