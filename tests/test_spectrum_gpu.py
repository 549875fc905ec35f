"""GPU parity of the spectrum display chain (glsCam/glsCam.cpp:196-202):
gls::magSpectrums / logMagSpectrums (glsBasicOperation.cpp:337-351),
gls::normalize (glsNormalize.cpp:52-75) and the fused fft + (log)magnitude,
against the restated oracle functions.

Bars: magnitudes within 2 ulp-scale relative error of the oracle's FP32
result (the GPU's sqrt/log may round differently in the last bit); NORM_MINMAX
bit-identical to the oracle (the same IEEE operations on the same min/max);
NORM_L1/L2 within 1e-6 relative (FP64 sums in another order); the fused forms
(fft + magnitude, and the three-launch glsCam view) bit-identical to the
separate calls; the whole glsCam chain, per element, within the bound the
FP32 FFT's own error at that bin implies (propagated through the log and the
min/max normalisation) of the oracle chain on the FP64 spectrum."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def rnd_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("shape", [(1, 1), (3, 5), (64, 64), (256, 512), (1024, 1024), (2048, 16)])
def test_mag_and_log_mag_vs_oracle(gls, shape):
    z = rnd_c(shape, sum(shape))
    z.flat[0] = 0  # exact zero: log(offset)
    m = gls.GlsMat(z)
    mag = gls.magSpectrums(m).download()
    assert mag.shape == shape and mag.dtype == np.float32
    assert np.allclose(mag, oracle.mag_spectrums(z), rtol=3e-7, atol=0)
    for off in (1.0, 1e-3, -0.5):
        lm = gls.logMagSpectrums(m, None, off).download()
        assert np.allclose(lm, oracle.log_mag_spectrums(z, off), rtol=2e-6, atol=2e-7)


def test_dst_reuse_and_aliasing(gls):
    z = rnd_c((32, 32), 1)
    m = gls.GlsMat(z)
    d = gls.GlsMat((32, 32), gls.CV_32FC1)
    out = gls.magSpectrums(m, d)
    assert out.handle == d.handle                        # matching dst is reused
    keep = m.share()                                     # a second handle on the spectrum
    same = gls.logMagSpectrums(m, m, 1.0)                # the glsCam call shape: dst is the src handle
    # dst = _dst: m now refers to a fresh CV_32FC1 plane; the spectrum's storage is untouched
    assert same.handle == m.handle and m.type() == gls.CV_32FC1 and m.handle != keep.handle
    assert np.array_equal(keep.download_complex(), z)
    assert np.allclose(m.download(), oracle.log_mag_spectrums(z, 1.0), rtol=2e-6, atol=2e-7)


@pytest.mark.parametrize("n", [1 << 10, 12345, 1 << 20])
def test_normalize_minmax_bit_exact(gls, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-4, 9, n).astype(np.float32).reshape(1, n)
    m = gls.GlsMat(x)
    for alpha, beta in ((0.0, 1.0), (2.0, -1.0), (10.0, 20.0)):
        y = gls.normalize(m, None, alpha, beta, gls.NORM_MINMAX).download()
        ref = oracle.normalize(x, alpha, beta, oracle.NORM_MINMAX)
        assert np.array_equal(y, ref)


@pytest.mark.parametrize("norm", [2, 4])
def test_normalize_l1_l2(gls, norm):
    rng = np.random.default_rng(norm)
    x = rng.standard_normal((512, 300)).astype(np.float32)
    y = gls.normalize(gls.GlsMat(x), None, 3.0, 0.0, norm).download()
    ref = oracle.normalize(x, 3.0, 0.0, norm)
    assert np.allclose(y, ref, rtol=1e-6, atol=0)
    got_norm = np.abs(y.astype(np.float64)).sum() if norm == 2 else np.linalg.norm(y.astype(np.float64))
    assert abs(got_norm - 3.0) < 1e-4


def test_normalize_in_place_constant_and_errors(gls):
    m = gls.GlsMat(np.full((16, 16), 7.0, np.float32))
    out = gls.normalize(m, m, 0.5, 1.0, gls.NORM_MINMAX)
    assert out.handle == m.handle and np.all(m.download() == 0.5)   # constant plane -> alpha
    with pytest.raises(gls.GlsError):
        gls.normalize(m, None, 1.0, 0.0, 1)                          # NORM_INF: not a normalize type
    with pytest.raises(gls.GlsError):
        gls.normalize(gls.GlsMat(rnd_c((4, 4), 2)), None, 1.0, 0.0, gls.NORM_L2)   # complex input
    with pytest.raises(gls.GlsError):
        gls.magSpectrums(gls.GlsMat(np.zeros((4, 4), np.float32)))                # real input


@pytest.mark.parametrize("shape,flags", [((256, 256), 64), ((512, 128), 0), ((64, 2048), 64 | 2),
                                         ((4096, 4096), 64), ((1, 4096), 64), ((2048, 1), 1)])
@pytest.mark.parametrize("log_mag", [False, True])
def test_fused_equals_fft_then_mag(gls, shape, flags, log_mag):
    rng = np.random.default_rng(shape[0] + shape[1] + flags)
    x = rng.uniform(0, 1, shape).astype(np.float32)
    src = gls.GlsMat(x)
    fused = gls.fftMagSpectrums(src, None, flags, log_mag, 1.0).download()
    spec = gls.fft(src, None, flags)
    sep = (gls.logMagSpectrums(spec, None, 1.0) if log_mag else gls.magSpectrums(spec)).download()
    assert np.array_equal(fused, sep)
    if shape[0] * shape[1] <= 1 << 20:
        ref = oracle.fft_f64(x.astype(np.complex128), flags)
        if log_mag:
            # d log(1 + |z|) <= |dz|: per element, the GPU's log magnitude is
            # the FP32 formula on the FP64 spectrum up to the FFT's own error at
            # that bin plus a few ulps of the formula's rounding
            ref32 = ref.astype(np.complex64)
            dz = np.abs(spec.download_complex().astype(np.complex128) - ref32.astype(np.complex128))
            want = oracle.log_mag_spectrums(ref32, 1.0)
            ulp = np.spacing(np.maximum(1.0, np.abs(want)).astype(np.float32)).astype(np.float64)
            assert np.all(np.abs(fused.astype(np.float64) - want) <= dz + 4 * ulp)
        else:
            want = np.abs(ref)
            assert np.linalg.norm(fused - want) / np.linalg.norm(want) <= 1e-5


@pytest.mark.parametrize("n", [256, 1024])
def test_glscam_fft_view_chain(gls, n):
    """fft(SHIFT) -> logMagSpectrums(1.0) -> normalize(0, 1, NORM_MINMAX), in place on one handle."""
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 1, (n, n)).astype(np.float32)
    frame = gls.GlsMat(x)
    frame = gls.fft(frame, frame, gls.GLS_FFT_SHIFT)
    frame = gls.logMagSpectrums(frame, frame, 1.0)
    frame = gls.normalize(frame, frame, 0.0, 1.0, gls.NORM_MINMAX)
    got = frame.download()
    spec = oracle.fft_f64(x.astype(np.complex128), oracle.SHIFT).astype(np.complex64)
    lm = oracle.log_mag_spectrums(spec, 1.0)
    ref = oracle.normalize(lm, 0.0, 1.0, oracle.NORM_MINMAX)
    # per element: the log magnitude moves by at most e = |dz| / (1 + |z|) plus a
    # few ulps (dz: the FFT's own error at the bin); the plane's min and max by
    # at most max(e) each; the normalisation divides by the span
    gpu_spec = gls.fft(gls.GlsMat(x), None, gls.GLS_FFT_SHIFT).download_complex()
    dz = np.abs(gpu_spec.astype(np.complex128) - spec.astype(np.complex128))
    e = dz / (1.0 + np.abs(spec.astype(np.complex128))) + 4 * np.spacing(np.maximum(np.abs(lm), 1.0)).astype(np.float64)
    span = float(lm.max()) - float(lm.min())
    bound = (e + 2 * e.max()) / span + 4 * float(np.spacing(np.float32(1)))
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound)
    assert abs(got.min()) < 1e-6 and abs(got.max() - 1.0) < 1e-6


@pytest.mark.parametrize("shape,flags", [((256, 256), 64), ((1024, 512), 64), ((64, 4096), 66), ((1, 2048), 64),
                                         ((4096, 4096), 64), ((1, 1024), 64), ((1, 512), 3), ((2048, 32), 67),
                                         ((1024, 4), 65), ((4096, 8), 0), ((8192, 8192), 64),
                                         ((4096, 4096), 0), ((8192, 4096), 0), ((4096, 8192), 3)])
def test_spectrum_view_equals_the_three_calls(gls, shape, flags):
    rng = np.random.default_rng(sum(shape))
    x = rng.uniform(0, 1, shape).astype(np.float32)
    src = gls.GlsMat(x)
    view = gls.spectrumView(src, None, flags, 1.0, 0.0, 1.0).download()
    sep = gls.normalize(gls.logMagSpectrums(gls.fft(src, None, flags), None, 1.0), None, 0.0, 1.0,
                        gls.NORM_MINMAX).download()
    # the same log magnitudes, the same min and max, the same affine map
    assert np.array_equal(view, sep)
    # alpha > beta and another offset
    v2 = gls.spectrumView(src, None, flags, 0.5, 3.0, -2.0).download()
    s2 = gls.normalize(gls.logMagSpectrums(gls.fft(src, None, flags), None, 0.5), None, 3.0, -2.0,
                       gls.NORM_MINMAX).download()
    assert np.array_equal(v2, s2)


@pytest.mark.parametrize("flags", [64, 0])
def test_spectrum_view_unaligned_destination(gls, flags):
    """A real 4096 x 4096 frame's view stores 16-byte groups of the expanded
    half plane; a destination that is 8-byte but not 16-byte aligned takes
    the whole-plane form instead, with the same bits."""
    n = 4096
    rng = np.random.default_rng(77 + flags)
    x = rng.uniform(0, 1, (n, n)).astype(np.float32)
    src = gls.GlsMat(x)
    ref = gls.spectrumView(src, None, flags, 1.0, 0.0, 1.0).download()
    # a plane of n*n + 4 floats (256-byte aligned), the view written from its third float on
    buf = gls.GlsMat((1, n * n + 4), gls.CV_32FC1)
    assert buf.dev_ptr % 16 == 0
    gls.spectrum_view_dev(src.dev_ptr, gls.CV_32FC1, buf.dev_ptr + 8, n, n, flags, 1.0, 0.0, 1.0)
    got = buf.download().reshape(-1)[2:2 + n * n].reshape(n, n)
    assert np.array_equal(got, ref)
