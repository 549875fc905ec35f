"""GPU parity of gls::fft (HIP path through the C ABI) against the FP64
oracle, the restated reference Stockham and the numpy golden vectors.

Bar: relative RMS error <= 1e-5 against FP64 for every result.
Cases follow unittest_glsFft.cpp:203-422 (16..2048 forward/inverse, SCALE,
SHIFT, 1 x N, N x 1) and extend to 4096^2, non-square shapes, real input and
in-place calls."""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rnd_c(shape, seed, dist="u"):
    rng = np.random.default_rng(seed)
    if dist == "u":
        return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def run_fft(gls, x, flags):
    src = gls.GlsMat(np.ascontiguousarray(x))
    dst = gls.fft(src, None, flags)
    return dst.download_complex()


FLAGS = [0, 1, 2, 3, 64, 65, 67]


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("flags", FLAGS)
def test_square_vs_fp64(gls, n, flags):
    x = rnd_c((n, n), n + flags)
    y = run_fft(gls, x, flags)
    ref = oracle.fft_f64(x, flags)
    err = oracle.rel_rms(y, ref)
    print(f"n={n} flags={flags} relrms={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("shape", [(1, 2), (2, 1), (2, 2), (1, 16), (16, 1), (1, 32), (64, 1), (1, 4096),
                                   (8192, 1), (2, 64), (64, 2), (4, 1024), (1024, 8), (64, 128), (128, 64),
                                   (256, 2048), (2048, 128), (16, 4096), (4096, 32), (1, 16384), (16384, 1),
                                   (16384, 4), (2, 16384), (8192, 8), (8192, 64), (1024, 4), (2048, 16)])
@pytest.mark.parametrize("flags", [0, 1, 3, 64, 67])
def test_shapes_vs_fp64(gls, shape, flags):
    x = rnd_c(shape, sum(shape) + flags)
    y = run_fft(gls, x, flags)
    assert oracle.rel_rms(y, oracle.fft_f64(x, flags)) <= TOL


@pytest.mark.parametrize("name", ["16x16", "32x32", "1x32", "64x1", "64x128", "8x4", "cfg1_256x256"])
def test_golden_vectors(gls, name):
    d = np.load(os.path.join(GOLD, f"fft_{name}.npz"))
    fl = {"fwd": 0, "inv": 1, "fwd_shift": 64, "inv_scale": 3}
    for key, flags in fl.items():
        if key in d:
            assert oracle.rel_rms(run_fft(gls, d["x"], flags), d[key]) <= TOL, key


def test_cfg1_256_uniform_roundtrip(gls):
    """Plain small: 256x256, U[-1,1], unscaled forward; fwd -> inv(SCALE)."""
    x = rnd_c((256, 256), 20261203)
    src = gls.GlsMat(x)
    spec = gls.fft(src)
    back = gls.fft(spec, None, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE)
    assert oracle.rel_rms(spec.download_complex(), oracle.fft_f64(x)) <= TOL
    assert oracle.rel_rms(back.download_complex(), x) <= TOL
    # the restated reference path agrees to the same bar
    assert oracle.rel_rms(spec.download_complex(), oracle.fft_r4_f32(x)) <= 2 * TOL


@pytest.mark.parametrize("n", [1024, 4096])
def test_cfg2_normal_fwd_inv(gls, n):
    """Plain on GPU at size: N(0,1) input, forward unscaled, inverse 1/(MN)."""
    x = rnd_c((n, n), 7 + n, "n")
    src = gls.GlsMat(x)
    spec = gls.fft(src)
    y = spec.download_complex()
    ref = oracle.fft_f64(x)
    e_fwd = oracle.rel_rms(y, ref)
    back = gls.fft(spec, None, gls.GLS_FFT_INVERSE | gls.GLS_FFT_SCALE).download_complex()
    e_inv = oracle.rel_rms(back, oracle.fft_f64(y, 3))
    e_rt = oracle.rel_rms(back, x)
    print(f"n={n} fwd={e_fwd:.3e} inv={e_inv:.3e} roundtrip={e_rt:.3e}")
    assert e_fwd <= TOL and e_inv <= TOL and e_rt <= TOL


def test_real_input_is_complex_with_zero_imag(gls):
    rng = np.random.default_rng(3)
    r = rng.uniform(-1, 1, (128, 256)).astype(np.float32)
    for flags in (0, 64, 3):
        y = gls.fft(gls.GlsMat(r), None, flags).download_complex()
        assert oracle.rel_rms(y, oracle.fft_f64(r.astype(np.complex128), flags)) <= TOL


def test_in_place(gls):
    x = rnd_c((512, 256), 11)
    m = gls.GlsMat(x)
    out = gls.fft(m, m, gls.GLS_FFT_SHIFT)
    assert out is m
    assert oracle.rel_rms(m.download_complex(), oracle.fft_f64(x, 64)) <= TOL


def test_dst_reuse_and_realloc(gls):
    x = rnd_c((64, 64), 12)
    src = gls.GlsMat(x)
    dst = gls.GlsMat((64, 64), gls.CV_32FC2)
    h = dst.handle
    gls.fft(src, dst)
    assert dst.handle == h                      # matching dst is reused
    small = gls.GlsMat((8, 8), gls.CV_32FC2)
    gls.fft(src, small)                         # wrong size -> reallocated (getDstMat)
    assert (small.rows, small.cols) == (64, 64)
    assert oracle.rel_rms(small.download_complex(), oracle.fft_f64(x)) <= TOL


def test_bad_inputs_raise(gls):
    with pytest.raises(gls.GlsError):
        gls.fft(gls.GlsMat(np.zeros((12, 16), np.float32)))          # not a power of two
    with pytest.raises(gls.GlsError):
        gls.fft(gls.GlsMat(np.zeros((16, 16), np.int32)))            # wrong depth
    with pytest.raises(gls.GlsError):
        gls.fft(gls.GlsMat(np.zeros((4, 32768), np.float32)))        # above 16384


def test_closed_forms_on_device(gls):
    m, n = 64, 128
    d = np.zeros((m, n), np.complex64)
    d[0, 0] = 1
    assert np.allclose(run_fft(gls, d, 0), 1.0, atol=1e-6)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    tone = np.exp(2j * np.pi * (5 * ii / m + 9 * jj / n)).astype(np.complex64)
    y = run_fft(gls, tone, 0)
    assert abs(y[5, 9] - m * n) < 1e-2 * m * n
    assert np.sort(np.abs(y).ravel())[-2] < 1e-3 * m * n


@pytest.mark.parametrize("length", [1024, 2048, 4096, 8192])
def test_long_columns_all_kernels_agree(gls, length):
    """1024- to 8192-point column passes run two columns per lane when the
    tile (8 columns; 4 at 8192 points) is full and the rows are 16-byte
    aligned; otherwise the one-column-per-lane kernels.  All must give the same
    bits, and FP64 within the bar."""
    C2 = gls.CV_32FC2
    x = rnd_c((length, 16), length + 41)
    src = gls.GlsMat(x)
    full = gls.GlsMat((length, 16), C2)
    col = np.fft.fft(x.astype(np.complex128), axis=0)
    for fl in (0, 1, 3, 64, 67):
        # all 16 columns: two tiles of the two-columns-per-lane kernel
        gls.fft1d_dev(src.dev_ptr, C2, full.dev_ptr, length, 16, 16, 1, fl)
        ref = full.download_complex()
        # a partial tile (12 columns; 6 at 8192 points) and 8 columns from an 8-byte offset: the other kernels
        k = 6 if length == 8192 else 12
        part = gls.GlsMat((length, 16), C2)
        gls.fft1d_dev(src.dev_ptr, C2, part.dev_ptr, length, k, 16, 1, fl)
        assert np.array_equal(part.download_complex()[:, :k].view(np.uint32), ref[:, :k].view(np.uint32)), fl
        off = gls.GlsMat((length, 16), C2)
        gls.fft1d_dev(src.dev_ptr + 8, C2, off.dev_ptr + 8, length, 8, 16, 1, fl)
        assert np.array_equal(off.download_complex()[:, 1:9].view(np.uint32), ref[:, 1:9].view(np.uint32)), fl
        if fl == 0:
            assert oracle.rel_rms(ref, col) <= TOL


@pytest.mark.parametrize("rows,cols,flags", [(4096, 4, 0), (4096, 8, 3), (4096, 64, 64), (4096, 16, 67),
                                            (4096, 4096, 1), (2048, 8, 67), (2048, 2048, 0), (1024, 32, 65),
                                            (1024, 1024, 3)])
def test_long_columns_plain_equals_plan(gls, rows, cols, flags):
    x = rnd_c((rows, cols), rows + cols + flags)
    y = run_fft(gls, x, flags)
    src, dst = gls.GlsMat(x), gls.GlsMat((rows, cols), gls.CV_32FC2)
    plan = gls.Plan.fft_batch([gls.FftDesc(src.dev_ptr, dst.dev_ptr, rows, cols, gls.CV_32FC2, flags)], 0,
                              keep=(src, dst))
    plan.execute()
    plan.check()
    assert np.array_equal(dst.download_complex().view(np.uint32), y.view(np.uint32))
    if rows * cols <= 1 << 20:
        assert oracle.rel_rms(y, oracle.fft_f64(x.astype(np.complex128), flags)) <= TOL
    # in place through the same kernels
    m = gls.GlsMat(x)
    assert np.array_equal(gls.fft(m, m, flags).download_complex().view(np.uint32), y.view(np.uint32))


def test_random_shapes_flags_inputs_vs_fp64(gls):
    """Seeded sweep over the dispatch space: power-of-two rows and cols 1..4096
    (at most 2^20 elements), every flag combination, real or complex input, in
    place or not -- whichever row and column kernels the shape selects (wave
    exchange, pair tiles, half-LDS tiles, generic tiles), each result within
    the FP64 bar."""
    rng = np.random.default_rng(2026)
    cases = 0
    while cases < 60:
        r, c = 1 << int(rng.integers(0, 13)), 1 << int(rng.integers(0, 13))
        if r * c > 1 << 20 or r * c < 2:
            continue
        cases += 1
        flags = int(rng.choice([0, 1, 2, 3, 64, 65, 66, 67]))
        real = rng.random() < 0.25
        if real:
            x = rng.uniform(-1, 1, (r, c)).astype(np.float32)
            y = gls.fft(gls.GlsMat(x), None, flags).download_complex()
            xc = x.astype(np.complex128)
        else:
            x = rnd_c((r, c), cases)
            m = gls.GlsMat(x)
            y = (gls.fft(m, m, flags) if rng.random() < 0.3 else gls.fft(m, None, flags)).download_complex()
            xc = x.astype(np.complex128)
        err = oracle.rel_rms(y, oracle.fft_f64(xc, flags))
        assert err <= TOL, (r, c, flags, real, err)


@pytest.mark.parametrize("cols,flags", [(2048, 0), (2048, 1), (2048, 3), (2048, 64), (2048, 67), (4096, 2)])
def test_decimated_columns_8192_rows_vs_fp64(gls, cols, flags):
    """Transforms of 8192 rows split the column transform by row parity (two
    4096-point column passes into scratch, the last radix-2 stage in the row
    pass).  Every flag set within the FP64 bar."""
    x = rnd_c((8192, cols), 8192 + cols + flags)
    y = run_fft(gls, x, flags)  # gls.fft(src, None, ...) writes a new dst: out of place
    assert oracle.rel_rms(y, oracle.fft_f64(x, flags)) <= TOL


@pytest.mark.parametrize("cols", [4096, 8192])
@pytest.mark.parametrize("flags", [0, 67])
def test_decimated_columns_in_place_equals_out_of_place(gls, cols, flags):
    """In place (src == dst) and out of place run the same passes (the column
    halves go to scratch either way): the same bits, repeatable; and the even
    output rows agree with the row-column path on 4096 rows: X[2k] over 8192
    rows is the 4096-row DFT of x[r] + x[r + 4096]."""
    x = rnd_c((8192, cols), 97 + cols + flags, "n")
    src = gls.GlsMat(x)
    a = gls.fft(src, None, flags).download_complex()
    b = gls.fft(src, None, flags).download_complex()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    m = gls.GlsMat(x)
    gls.fft(m, m, flags)
    assert np.array_equal(a.view(np.uint32), m.download_complex().view(np.uint32))
    if flags == 0:
        folded = gls.fft(gls.GlsMat(x[:4096] + x[4096:]), None, 0).download_complex()
        assert oracle.rel_rms(a[0::2], folded) <= 2 * TOL
    del x, a, b


@pytest.mark.parametrize("rows,cols", [(16, 16), (16, 32), (32, 16), (16, 64), (64, 16), (32, 32), (64, 64)])
@pytest.mark.parametrize("flags", [0, 1, 3, 64, 67])
@pytest.mark.parametrize("real", [False, True])
def test_small_frames_one_launch_equal_two_passes(gls, rows, cols, flags, real):
    """Frames of at most 1024 points (sides 16..64) run the whole 2D FFT in
    one workgroup: the same bits as a row pass followed by a column pass (two
    1D calls over the rows, then the columns, with the same flags), in place
    too; 64 x 64 takes the two passes in both forms."""
    rng = np.random.default_rng(rows * 1000 + cols + flags)
    if real:
        x = rng.uniform(-1, 1, (rows, cols)).astype(np.float32)
        typ = gls.CV_32FC1
    else:
        x = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))).astype(np.complex64)
        typ = gls.CV_32FC2
    src = gls.GlsMat(x)
    one = gls.GlsMat((rows, cols), gls.CV_32FC2)
    gls.fft_dev(src.dev_ptr, typ, one.dev_ptr, rows, cols, flags)
    two = gls.GlsMat((rows, cols), gls.CV_32FC2)
    gls.fft1d_dev(src.dev_ptr, typ, two.dev_ptr, cols, rows, 1, cols, flags)                  # the rows
    gls.fft1d_dev(two.dev_ptr, gls.CV_32FC2, two.dev_ptr, rows, cols, cols, 1, flags)         # the columns
    a, b = one.download_complex(), two.download_complex()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert oracle.rel_rms(a, oracle.fft_f64(x.astype(np.complex64), flags)) <= TOL
    if not real:
        gls.fft_dev(src.dev_ptr, typ, src.dev_ptr, rows, cols, flags)                         # in place
        assert np.array_equal(src.download_complex().view(np.uint32), a.view(np.uint32))
