"""GPU parity of the phase-correlation chain's standalone operations
(glscv_amd/csrc/chain_ops.hip): gls::multiply, mulSpectrums (with and without
conj), mulSpectrumsPhaseOnly, split, merge and minMaxLoc, each against the
oracle's restatement (oracle/oracle.c) or numpy's single IEEE operation, bit
for bit, through the buffer API and the raw-pointer forms (aligned and
unaligned pointers, odd counts); and the reference's chain composed of them
as glsPhaseCorrelate.cpp:174-210 writes it, whose integer peak must equal the
fused gls::phaseCorrelate's, the oracle chain's and the ground truth."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

FLT_MAX = np.float32(np.finfo(np.float32).max)


def same_bits(x, y):
    """Bitwise equality, with any NaN equal to any NaN (payloads differ
    between the host and the device)."""
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    both_nan = np.isnan(x) & np.isnan(y)
    return bool(np.all(both_nan | (x.view(np.uint32) == y.view(np.uint32))))


def rnd_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def c2(z):
    """complex64 (rows, cols) -> float32 (rows, cols, 2)"""
    return np.ascontiguousarray(z).view(np.float32).reshape(z.shape[0], z.shape[1], 2)


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (64, 64), (333, 517)])
@pytest.mark.parametrize("cn", [1, 2])
def test_multiply_bitwise(gls, shape, cn):
    rng = np.random.default_rng(shape[0] * 7 + cn)
    full = shape + ((cn,) if cn == 2 else ())
    a = rng.standard_normal(full).astype(np.float32)
    b = rng.standard_normal(full).astype(np.float32)
    ga, gb = gls.GlsMat(a), gls.GlsMat(b)
    out = gls.multiply(ga, gb)
    assert same_bits(out.download(), a * b)
    # getDstMat: a fitting dst is reused, a dst that is a source is replaced
    keep = gls.GlsMat(full[:2], ga.type())
    h = keep.handle
    gls.multiply(ga, gb, keep)
    assert keep.handle == h and same_bits(keep.download(), a * b)
    alias = ga.share()
    gls.multiply(alias, gb, alias)
    assert alias.handle != ga.handle and same_bits(alias.download(), a * b)
    assert same_bits(ga.download(), a)


def test_multiply_dev_unaligned_and_in_place(gls):
    rng = np.random.default_rng(3)
    n = 1001
    a = rng.standard_normal(n + 3).astype(np.float32)
    b = rng.standard_normal(n + 3).astype(np.float32)
    ga, gb = gls.GlsMat(a.reshape(1, -1)), gls.GlsMat(b.reshape(1, -1))
    gd = gls.GlsMat((1, n + 3), gls.CV_32FC1)
    # 4-byte offsets: the scalar path
    gls.multiply_dev(ga.dev_ptr + 4, gb.dev_ptr + 4, gd.dev_ptr + 4, n)
    gls.stream_sync()
    assert same_bits(gd.download()[0, 1:n + 1], a[1:n + 1] * b[1:n + 1])
    # in place (dst == src0), aligned: the vector path with a tail
    gls.multiply_dev(ga.dev_ptr, gb.dev_ptr, ga.dev_ptr, n + 3)
    gls.stream_sync()
    assert same_bits(ga.download()[0], a * b)


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (16, 16), (255, 129)])
@pytest.mark.parametrize("conj", [False, True])
def test_mul_spectrums_bitwise(gls, shape, conj):
    a, b = rnd_c(shape, 1), rnd_c(shape, 2)
    out = gls.mulSpectrums(gls.GlsMat(a), gls.GlsMat(b), None, conj).download_complex()
    assert same_bits(out.view(np.float32), oracle.mul_spectrums(a, b, conj).view(np.float32))


def phase_only_inputs():
    a, b = rnd_c((40, 50), 5), rnd_c((40, 50), 6)
    a[0, :5] = 0
    b[1, :5] = 0
    a[2, :4] = np.array([1e-30, 1e-20, 1e-19 + 1e-19j, 3e-23], np.complex64)      # |c| around the guard
    b[2, :4] = np.array([1e-10, 1e-15, 1e-20, 1.0], np.complex64)
    a[3, :3] = np.array([np.inf, np.nan, 1e20 + 1e20j], np.complex64)
    a[4, :2] = np.array([3e38, -3e38j], np.complex64)                                 # |c|^2 overflows
    b[4, :2] = np.array([2.0, 2.0], np.complex64)
    return a, b


def test_phase_only_bitwise_with_guard(gls):
    a, b = phase_only_inputs()
    out = gls.mulSpectrumsPhaseOnly(gls.GlsMat(a), gls.GlsMat(b)).download_complex()
    ref = oracle.poc(a, b)
    assert same_bits(out.view(np.float32), ref.view(np.float32))
    assert np.all(out[0, :5] == 0) and np.all(out[1, :5] == 0)
    # unit magnitude where |c|^2 is a normal float (rows 5 on; the planted
    # rows follow the shader's formula into underflow and overflow, bit for
    # bit with the oracle above)
    assert np.allclose(np.abs(out[5:]), 1.0, atol=1e-6)


def test_spectrum_dev_modes_unaligned(gls):
    """8-byte aligned but not 16-byte aligned pointers and an odd count: the
    scalar path, same bits as the vector path."""
    n = 777
    a, b = rnd_c((1, n + 1), 8), rnd_c((1, n + 1), 9)
    ga, gb = gls.GlsMat(a), gls.GlsMat(b)
    gd = gls.GlsMat((1, n + 1), gls.CV_32FC2)
    for mode, ref in ((gls.SPECTRUM_MUL, oracle.mul_spectrums(a[:, 1:], b[:, 1:])),
                      (gls.SPECTRUM_MUL_CONJ, oracle.mul_spectrums(a[:, 1:], b[:, 1:], True)),
                      (gls.SPECTRUM_PHASE_ONLY, oracle.poc(a[:, 1:], b[:, 1:]))):
        gls.mul_spectrums_dev(ga.dev_ptr + 8, gb.dev_ptr + 8, gd.dev_ptr + 8, n, mode)
        gls.stream_sync()
        assert same_bits(gd.download_complex()[:, 1:].view(np.float32), ref.view(np.float32)), mode
    with pytest.raises(gls.GlsError):
        gls.mul_spectrums_dev(ga.dev_ptr + 4, gb.dev_ptr, gd.dev_ptr, n, gls.SPECTRUM_MUL)
    with pytest.raises(gls.GlsError):
        gls.mul_spectrums_dev(ga.dev_ptr, gb.dev_ptr, gd.dev_ptr, n, 3)


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (128, 64), (301, 257)])
def test_split_merge_round_trip(gls, shape):
    z = rnd_c(shape, 10)
    g = gls.GlsMat(z)
    re, im = gls.split(g)
    assert re.type() == gls.CV_32FC1 and im.type() == gls.CV_32FC1
    assert same_bits(re.download(), z.real) and same_bits(im.download(), z.imag)
    back = gls.merge([re, im])
    assert back.type() == gls.CV_32FC2 and same_bits(back.download(), c2(z))
    keep = gls.GlsMat(shape, gls.CV_32FC2)
    h = keep.handle
    gls.merge([im, re], keep)                                   # a fitting dst is reused
    assert keep.handle == h and same_bits(keep.download(), c2(z.imag + 1j * z.real))
    one, = gls.split(re)                                        # one channel: a copy
    assert one.handle != re.handle and same_bits(one.download(), z.real)
    assert same_bits(gls.merge([re]).download(), z.real)


def test_split_merge_dev_forms(gls):
    n = 1023
    z = rnd_c((1, n + 1), 11)
    g = gls.GlsMat(z)
    p0 = gls.GlsMat((1, n + 1), gls.CV_32FC1)
    p1 = gls.GlsMat((1, n + 1), gls.CV_32FC1)
    # channel 0 only (phaseCorrelate keeps the real plane), unaligned source
    gls.split_dev(g.dev_ptr + 8, p0.dev_ptr + 4, None, n)
    gls.stream_sync()
    assert same_bits(p0.download()[0, 1:], z.real[0, 1:])
    gls.split_dev(g.dev_ptr, p0.dev_ptr, p1.dev_ptr, n + 1)
    out = gls.GlsMat((1, n + 1), gls.CV_32FC2)
    gls.merge_dev(p0.dev_ptr + 4, p1.dev_ptr + 4, out.dev_ptr + 8, n)
    gls.stream_sync()
    assert same_bits(out.download()[0, 1:], c2(z)[0, 1:])


def minmax_cases():
    rng = np.random.default_rng(42)
    cases = []
    for shape in ((1, 1), (1, 9), (9, 1), (3, 5), (37, 53), (1, 100003), (1000, 1234)):
        cases.append(rng.standard_normal(shape).astype(np.float32))
        cases.append(rng.integers(-3, 4, shape).astype(np.float32))
    x = rng.integers(-2, 3, (300, 257)).astype(np.float32)
    x[rng.random(x.shape) < 0.3] = np.nan
    cases.append(x)
    y = rng.standard_normal((64, 64)).astype(np.float32)
    y[2, 3], y[5, 1], y[0, 7] = np.inf, -np.inf, np.nan
    cases.append(y)
    for fill in (np.nan, -np.inf, np.inf, -FLT_MAX, FLT_MAX, 0.0):
        cases.append(np.full((33, 65), fill, np.float32))
    z = np.zeros((5, 7), np.float32)
    z[3, 2], z[1, 6] = -FLT_MAX, FLT_MAX
    cases.append(z)
    # equal extremes far apart: the first in row-major order wins across blocks
    big = rng.standard_normal((2048, 2048)).astype(np.float32)
    for (r, c) in ((2047, 2047), (1500, 3), (700, 2000), (1500, 2)):
        big[r, c] = 10.0
        big[r, (c + 1) % 2048] = -10.0
    cases.append(big)
    cases.append(rng.standard_normal((4096, 4096)).astype(np.float32))
    two = rng.standard_normal((511, 257, 2)).astype(np.float32)
    two[..., 1] *= 100.0                                        # channel 1 is ignored
    cases.append(two)
    return cases


def test_min_max_loc_vs_oracle(gls):
    for k, v in enumerate(minmax_cases()):
        got = gls.minMaxLoc(gls.GlsMat(v))
        want = oracle.min_max_loc(v)
        assert got == want, (k, v.shape, got, want)


def test_min_max_loc_dev_async_unaligned(gls):
    rng = np.random.default_rng(7)
    rows, cols = 129, 77
    v = rng.integers(-5, 6, (rows * cols + 1,)).astype(np.float32)
    g = gls.GlsMat(v.reshape(1, -1))
    res = gls.GlsMat((1, 6), gls.CV_32SC1)                      # a glsx_minmax_result in device memory
    gls.min_max_loc_dev(g.dev_ptr + 4, gls.CV_32FC1, rows, cols, res.dev_ptr)
    gls.stream_sync()
    r = res.download().reshape(-1)
    want = oracle.min_max_loc(v[1:].reshape(rows, cols))
    assert (float(r[:1].view(np.float32)[0]), float(r[1:2].view(np.float32)[0]),
            (int(r[2]), int(r[3])), (int(r[4]), int(r[5]))) == want


def test_argument_errors(gls):
    a = gls.GlsMat((8, 8), gls.CV_32FC1)
    b = gls.GlsMat((8, 4), gls.CV_32FC1)
    c = gls.GlsMat((8, 8), gls.CV_32FC2)
    with pytest.raises(gls.GlsError):
        gls.multiply(a, b)                                      # sizes differ
    with pytest.raises(gls.GlsError):
        gls.multiply(a, c)                                      # types differ
    with pytest.raises(gls.GlsError):
        gls.mulSpectrums(a, a)                                  # not a spectrum
    with pytest.raises(gls.GlsError):
        gls.merge([a, b])                                       # planes differ
    with pytest.raises(gls.GlsError):
        gls.merge([a, a, a])                                    # three planes
    ints = gls.GlsMat((8, 8), gls.CV_32SC1)
    with pytest.raises(gls.GlsError):
        gls.minMaxLoc(ints)                                     # float images only


def test_dev_forms_reject_partial_overlap(gls):
    """Raw-pointer forms: a destination that is a source exactly runs in place,
    one that overlaps a source partly (or split/merge planes that overlap the
    image) is refused on the host before any launch."""
    buf = gls.GlsMat((1, 4096), gls.CV_32FC2)
    p = buf.dev_ptr
    n = 1000
    gls.multiply_dev(p, p + 8000, p, n)                         # dst == src0: in place
    gls.mul_spectrums_dev(p, p + 8000, p, n, gls.SPECTRUM_MUL)
    gls.stream_sync()
    for call in (lambda: gls.multiply_dev(p, p + 8000, p + 16, n),
                 lambda: gls.mul_spectrums_dev(p, p + 8000, p + 8, n, gls.SPECTRUM_MUL_CONJ),
                 lambda: gls.split_dev(p, p + 4000, None, n),
                 lambda: gls.split_dev(p, p + 16000, p + 18000, n),
                 lambda: gls.merge_dev(p + 16000, p + 20000, p + 12000, n)):
        with pytest.raises(gls.GlsError) as e:
            call()
        assert e.value.status == -1 and "overlap" in str(e.value)


def box5(img):
    acc = np.zeros_like(img, dtype=np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            acc += np.roll(img, (dy, dx), axis=(0, 1))
    return (acc / 25.0).astype(np.float32)


@pytest.mark.parametrize("n,seed,window", [(64, 1, False), (256, 2, False), (512, 3, True), (1024, 4, False),
                                           (2048, 5, True)])
def test_reference_chain_by_operations(gls, n, seed, window):
    """glsPhaseCorrelate.cpp:174-210, one operation per call: window multiply,
    fft, fft, mulSpectrumsPhaseOnly, fft(INVERSE | SHIFT) in place, split,
    minMaxLoc.  Its peak equals the ground truth, the oracle chain's and the
    fused gls::phaseCorrelate's; its surface is the oracle's within the FP32
    error of the two FFTs."""
    rng = np.random.default_rng(seed)
    img1 = box5(rng.integers(0, 256, (n, n), dtype=np.uint8).astype(np.float32) / 255.0)
    dx, dy = int(rng.integers(-n // 4, n // 4)), int(rng.integers(-n // 4, n // 4))
    img2 = (np.roll(img1, (dy, dx), axis=(0, 1)) + rng.normal(0, 0.05, (n, n))).astype(np.float32)
    win = None
    if window:
        w1 = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)
        win = np.outer(w1, w1).astype(np.float32)
    padded1, padded2 = gls.GlsMat(img1), gls.GlsMat(img2)
    if win is not None:
        gwin = gls.GlsMat(win)
        padded1 = gls.multiply(gwin, padded1, padded1)
        padded2 = gls.multiply(gwin, padded2, padded2)
    fft1 = gls.fft(padded1)
    fft2 = gls.fft(padded2)
    C = gls.mulSpectrumsPhaseOnly(fft1, fft2)
    gls.fft(C, C, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SHIFT)
    C = gls.split(C)[0]
    _, _, _, peak = gls.minMaxLoc(C)
    ref, surf = oracle.phase_correlate(img1, img2, win, want_surface=True)
    _, _, fused_peak = gls.phaseCorrelate(gls.GlsMat(img1), gls.GlsMat(img2),
                                          gls.GlsMat(win) if win is not None else None)
    print(f"n={n} truth=({n // 2 - dx},{n // 2 - dy}) ops={peak} fused={fused_peak} oracle=({ref.peak_x},{ref.peak_y})")
    assert peak == (n // 2 - dx, n // 2 - dy)
    assert peak == (ref.peak_x, ref.peak_y) == tuple(fused_peak)
    # the surface against the same chain in FP64: the phase-only product gives
    # every bin unit weight, so bins of small |c| carry their FP32 phase error
    # at full weight (1e-5..1e-4 rel-RMS for either FP32 chain); the
    # operations' surface must be as close as the oracle's FP32 restatement
    # of the reference chain, within a factor 2
    w = win if win is not None else np.ones((n, n), np.float32)
    f1 = np.fft.fft2((img1 * w).astype(np.float64))
    f2 = np.fft.fft2((img2 * w).astype(np.float64))
    c = f1 * np.conj(f2)
    m = np.abs(c)
    p = np.where(m > 1e-30, c / np.where(m > 0, m, 1), 0)
    surf64 = np.fft.fftshift(np.real(np.fft.ifft2(p)) * n * n)
    err_ops = oracle.rel_rms(C.download(), surf64)
    err_orc = oracle.rel_rms(surf, surf64)
    print(f"surface rel-RMS vs FP64: operations {err_ops:.2e}, oracle {err_orc:.2e}")
    assert err_ops <= 2 * err_orc and err_ops < 1e-3
