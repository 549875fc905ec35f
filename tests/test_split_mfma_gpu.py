"""GPU parity of the split-precision MFMA FFT (GLS_FFT_SPLIT_MFMA): radix-16
stages as 16x16 DFT-matrix products on the matrix cores with a 3-term BF16
split and FP32 accumulation.

Bar: relative RMS error <= 1e-5 against FP64 for every result, the same bar
as the FP32 kernels; per dimension the library's own FP64 check must hold
1e-5/sqrt(2).  The FP32 kernel's error is printed beside the MFMA error."""
import numpy as np
import pytest

import oracle
from glscv_amd.workloads import tones_2d

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rnd_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


def run(gls, x, flags):
    src = gls.GlsMat(np.ascontiguousarray(x))
    return gls.fft(src, None, flags).download_complex()


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fp64_check_per_length(gls, n):
    err, enabled = gls.split_mfma_check(n)
    print(f"len={n} split-MFMA relrms vs FP64 DFT = {err:.3e} enabled={enabled}")
    # a real rounding error, not a copy of an exact result, and within the bound
    assert 1e-9 < err <= gls.SPLIT_MFMA_TOL
    assert enabled
    # the kept result is returned again
    assert gls.split_mfma_check(n) == (err, enabled)


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_short_lengths_not_enabled(gls, n):
    # no radix-16 stage: nothing to run on the matrix cores
    assert gls.split_mfma_check(n) == (0.0, False)


def test_check_rejects_bad_length(gls):
    with pytest.raises(gls.GlsError):
        gls.split_mfma_check(48)


@pytest.mark.parametrize("seed", [0, 1])
def test_tones_2048_side_by_side(gls, seed):
    """8 random-frequency tones (amplitudes 1e-3..1e3) + N(0, 1e-2) noise."""
    x = tones_2d(2048, 2048, seed)
    ref = np.fft.fft2(x.astype(np.complex128))
    e32 = oracle.rel_rms(run(gls, x, 0), ref)
    emf = oracle.rel_rms(run(gls, x, gls.GLS_FFT_SPLIT_MFMA), ref)
    print(f"2048^2 tones seed={seed}: relrms FP32 {e32:.3e}  split-MFMA {emf:.3e}")
    assert e32 <= TOL
    assert emf <= TOL


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("flags", [0, 1, 3, 64, 67])
def test_square_vs_fp64(gls, n, flags):
    x = rnd_c((n, n), 7 * n + flags)
    y = run(gls, x, flags | gls.GLS_FFT_SPLIT_MFMA)
    ref = oracle.fft_f64(x, flags)
    err = oracle.rel_rms(y, ref)
    print(f"n={n} flags={flags} split-MFMA relrms={err:.3e}")
    assert err <= TOL


# shapes with short and long dimensions, partial waves (few transforms per
# workgroup) and the 1D forms
@pytest.mark.parametrize("shape", [(16, 16), (16, 32), (32, 16), (4, 64), (64, 4), (2, 2048), (2048, 2),
                                   (1, 16), (16, 1), (1, 4096), (8192, 1), (1, 16384), (16384, 1),
                                   (16, 8192), (8192, 16), (64, 2048), (512, 32), (8, 16384)])
@pytest.mark.parametrize("flags", [0, 3, 67])
def test_shapes_vs_fp64(gls, shape, flags):
    x = rnd_c(shape, sum(shape) + flags)
    y = run(gls, x, flags | gls.GLS_FFT_SPLIT_MFMA)
    assert oracle.rel_rms(y, oracle.fft_f64(x, flags)) <= TOL


def test_real_input(gls):
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (256, 512)).astype(np.float32)
    y = gls.fft(gls.GlsMat(x), None, gls.GLS_FFT_SPLIT_MFMA).download_complex()
    assert oracle.rel_rms(y, np.fft.fft2(x.astype(np.float64))) <= TOL


def test_roundtrip(gls):
    x = rnd_c((1024, 1024), 11)
    m = gls.GlsMat(x)
    gls.fft(m, m, gls.GLS_FFT_SPLIT_MFMA)
    gls.fft(m, m, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE | gls.GLS_FFT_SPLIT_MFMA)
    assert oracle.rel_rms(m.download_complex(), x) <= TOL


def test_differs_from_fp32_kernel(gls):
    # the flag really selects another kernel: bits differ from the FP32 path
    x = rnd_c((256, 256), 3)
    a = run(gls, x, 0)
    b = run(gls, x, gls.GLS_FFT_SPLIT_MFMA)
    assert not np.array_equal(a, b)
    assert oracle.rel_rms(b, a) <= 2 * TOL


def test_plans_reject_the_flag(gls):
    m = gls.GlsMat(rnd_c((64, 64), 1))
    with pytest.raises(gls.GlsError):
        gls.Plan.fft_batch([gls.FftDesc(m.dev_ptr, m.dev_ptr, 64, 64, gls.CV_32FC2, gls.GLS_FFT_SPLIT_MFMA)])


@pytest.mark.parametrize("n", [2048, 4096, 8192])
@pytest.mark.parametrize("flags", [0, 3, 64])
def test_row_kernel_matches_lds_tile(gls, n, flags):
    """Two rows of n points run the plain row kernel with the matrix-core
    butterfly (register layout, lane swaps around the products); the two
    columns of the transposed array run the column tile that passes its data
    through LDS in the products' layout.  Same twiddles, products and sums:
    the same bits."""
    x = rnd_c((2, n), 13 * n + flags)
    f = flags | gls.GLS_FFT_SPLIT_MFMA
    src_r, dst_r = gls.GlsMat(x), gls.GlsMat(np.zeros_like(x))
    gls.fft1d_dev(src_r.dev_ptr, gls.CV_32FC2, dst_r.dev_ptr, n, 2, 1, n, f)
    xt = np.ascontiguousarray(x.T)
    src_c, dst_c = gls.GlsMat(xt), gls.GlsMat(np.zeros_like(xt))
    gls.fft1d_dev(src_c.dev_ptr, gls.CV_32FC2, dst_c.dev_ptr, n, 2, 2, 1, f)
    rows, cols = dst_r.download_complex(), dst_c.download_complex()
    assert np.array_equal(rows, cols.T)
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    if flags & gls.GLS_FFT_INVERSE:
        ref = np.fft.ifft(x.astype(np.complex128), axis=1) * (1 if flags & gls.GLS_FFT_SCALE else n)
    if flags & gls.GLS_FFT_SHIFT:
        ref = np.roll(ref, n // 2, axis=1)
    assert oracle.rel_rms(rows, ref) <= TOL
