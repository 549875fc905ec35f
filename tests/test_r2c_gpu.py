"""gls::fft of a real image (CV_32FC1).  From 4096 x 4096 up it runs through
half-size spectra (glscv_amd/csrc/fft_r2c.hip: 20 instead of 28 bytes per
pixel); below, where that measured slower, through the complex transform of
the image with imaginary part 0.  The reference merges the image with an empty
imaginary plane and runs the complex transform (glsCV/glsFft.cpp:683-689).

Parity: every result within the FP64 bar (rel-RMS <= 1e-5, the bar of the
complex path), at every power-of-two row and column length and flag
combination.  Properties that do not depend on size, on the half-size path:
the output is Hermitian exactly, Y[-r][-c] == conj(Y[r][c]) (every mirror
value is stored as the conjugate of the value it mirrors), it is the same on
every run, the inverse is the conjugate of the forward transform bit for bit,
and it agrees with the complex transform of the same image within the FP64
bar of both.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5
FLAG_SETS = (0, 1, 2, 3, 64, 65, 66, 67)


def _real(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(np.float32)


def _check(gls, x, flags):
    y = gls.fft(gls.GlsMat(x), None, flags).download_complex()
    err = oracle.rel_rms(y, oracle.fft_f64(x.astype(np.complex128), flags))
    assert err <= TOL, (x.shape, flags, err)
    return y


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048])
def test_square_every_flag_vs_fp64(gls, n):
    x = _real((n, n), n)
    for flags in FLAG_SETS:
        _check(gls, x, flags)


@pytest.mark.parametrize("shape", [(2, 2), (2, 64), (64, 2), (4, 4096), (4096, 4), (8, 16384), (16384, 8),
                                   (32, 1024), (1024, 32), (256, 2048), (2048, 256), (16, 8192), (8192, 16),
                                   (128, 512), (512, 16)])
def test_rectangular_vs_fp64(gls, shape):
    x = _real(shape, shape[0] * 7 + shape[1])
    for flags in (0, 67):
        _check(gls, x, flags)


def test_half_size_path_every_flag_vs_fp64(gls):
    x = _real((4096, 4096), 1)
    for flags in FLAG_SETS:
        _check(gls, x, flags)


@pytest.mark.parametrize("shape", [(4096, 8192), (8192, 4096)])
def test_half_size_path_rectangular_vs_fp64(gls, shape):
    x = _real(shape, 2)
    for flags in (0, 3, 64):
        _check(gls, x, flags)


def _hermitian_exact(y):
    """Y[-r][-c] == conj(Y[r][c]) exactly (as values: the imaginary part of
    a point that is its own mirror is +0.0, its conjugate -0.0)."""
    m, n = y.shape
    rr = (-np.arange(m)) % m
    cc = (-np.arange(n)) % n
    mirror = y[rr][:, cc]
    bad = np.argwhere((mirror.real != y.real) | (-mirror.imag != y.imag))
    assert bad.size == 0, f"{len(bad)} points not mirrored exactly, first {bad[:5].tolist()}"
    return True


@pytest.mark.parametrize("shape", [(8192, 8192), (4096, 4096), (4096, 16384), (16384, 4096)])
def test_full_size_hermitian_and_repeatable(gls, shape):
    """At sizes the FP64 oracle would take long on, the size-independent
    properties: the spectrum of a real image is exactly Hermitian (bit for bit),
    the same on every run, and within the FP64 bar of the complex transform of
    the same data."""
    x = _real(shape, 5)
    src = gls.GlsMat(x)
    y1 = gls.fft(src).download_complex()
    y2 = gls.fft(src).download_complex()
    assert np.array_equal(y1.view(np.uint64), y2.view(np.uint64))
    assert _hermitian_exact(y1)
    yc = gls.fft(gls.GlsMat(x.astype(np.complex64))).download_complex()
    assert oracle.rel_rms(y1, yc) <= 2 * TOL


def test_inverse_is_conjugate_of_forward_on_real_input(gls):
    x = _real((4096, 4096), 9)
    src = gls.GlsMat(x)
    f = gls.fft(src, None, 0).download_complex()
    i = gls.fft(src, None, gls.GLS_FFT_INVERSE).download_complex()
    assert np.array_equal(i.real.view(np.uint32), f.real.view(np.uint32))
    assert np.array_equal(i.imag.view(np.uint32), (-f.imag).view(np.uint32))


def test_real_frame_spectrum_view_vs_fp64(gls):
    """glsCam's view (glsCam.cpp:199-202) on a real frame runs the real-input
    column pass storing log magnitudes with (max, min) partials."""
    for n, shape in ((0, (512, 512)), (1, (4096, 4096)), (2, (4096, 8192)), (3, (8192, 64))):
        x = _real(shape, 30 + n)
        view = gls.spectrumView(gls.GlsMat(x), None, gls.GLS_FFT_SHIFT, 1.0, 0.0, 1.0).download()
        spec = np.fft.fftshift(np.fft.fft2(x.astype(np.float64)))
        lm = np.log(np.abs(spec) + 1.0)
        ref = (lm - lm.min()) / (lm.max() - lm.min())
        assert np.max(np.abs(view - ref)) < 1e-4, shape
        mag = gls.fftMagSpectrums(gls.GlsMat(x), None, 0, False).download()
        ref_mag = np.abs(np.fft.fft2(x.astype(np.float64)))
        assert np.max(np.abs(mag - ref_mag)) <= 1e-5 * np.abs(ref_mag).max() * 8, shape


def test_overlapping_source_rejected(gls):
    """A real source read while the half spectra are written must not share
    memory with the destination (checked on the host before any launch)."""
    buf = gls.GlsMat((64, 64), gls.CV_32FC2)
    with pytest.raises(gls.GlsError) as e:
        gls.fft_dev(buf.dev_ptr, gls.CV_32FC1, buf.dev_ptr, 64, 64, 0)
    assert e.value.status == -1 and "overlap" in str(e.value)
