"""GPU parity of gls::phaseCorrelate (device chain through the C ABI) against
the restated CPU chain (oracle) and the ground-truth shift.

Bar: integer peaks equal exactly (ground truth and CPU oracle); the
sub-pixel shift and the response agree with the oracle within the FP32 error
of the two FFT implementations.  Inputs follow the serving workload: uint8
noise / 255, 5x5 circular box blur, integer circular shift in [-N/4, N/4) plus
N(0, 0.05) noise (UnitTest_glsPhaseCorrelate.cpp:45-147 uses a sub-pixel
warpAffine shift instead, which needs OpenCV)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def box5(img):
    acc = np.zeros_like(img, dtype=np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            acc += np.roll(img, (dy, dx), axis=(0, 1))
    return (acc / 25.0).astype(np.float32)


def make_pair(n, seed, noise=0.05, m=None):
    m = m or n
    rng = np.random.default_rng(seed)
    base = box5(rng.integers(0, 256, (m, n), dtype=np.uint8).astype(np.float32) / 255.0)
    dx = int(rng.integers(-n // 4, n // 4))
    dy = int(rng.integers(-m // 4, m // 4))
    img2 = np.roll(base, (dy, dx), axis=(0, 1))
    if noise:
        img2 = (img2 + rng.normal(0, noise, img2.shape)).astype(np.float32)
    return base, img2.astype(np.float32), dx, dy


def hann(m, n):
    wy = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(m) / m)
    wx = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n) / n)
    return np.outer(wy, wx).astype(np.float32)


@pytest.mark.parametrize("n,seed", [(32, 1), (64, 2), (256, 3), (512, 4), (1024, 5), (2048, 6), (256, 7),
                                    (512, 8)])
def test_integer_peak_matches_truth_and_oracle(gls, n, seed):
    a, b, dx, dy = make_pair(n, seed)
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b))
    ref, _ = oracle.phase_correlate(a, b)
    print(f"n={n} truth=({dx},{dy}) gpu_peak=({px},{py}) cpu_peak=({ref.peak_x},{ref.peak_y}) "
          f"shift=({sx:.3f},{sy:.3f}) resp={resp:.4f}/{ref.response:.4f}")
    assert (px, py) == (n // 2 - dx, n // 2 - dy)
    assert (px, py) == (ref.peak_x, ref.peak_y)
    assert abs(sx - dx) < 0.5 and abs(sy - dy) < 0.5
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)
    assert resp == pytest.approx(ref.response, rel=1e-3, abs=1e-4)


def test_non_square_and_window(gls):
    a, b, dx, dy = make_pair(256, 9, m=128)
    w = hann(128, 256)
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b), gls.GlsMat(w))
    ref, _ = oracle.phase_correlate(a, b, w)
    assert (px, py) == (ref.peak_x, ref.peak_y) == (128 - dx, 64 - dy)
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)


@pytest.mark.parametrize("rows,cols,seed", [(4096, 16, 13), (8192, 8, 14), (1024, 4, 15), (2048, 64, 16)])
def test_tall_images_surface_tiles(gls, rows, cols, seed):
    """Tall images: the surface pass runs 8-column tiles with two columns per
    lane at 1024 and 2048 rows, storing the real plane with argmax partials,
    and the one-column tiles at 4096 and 8192 rows and for 4 columns (not a
    whole tile)."""
    a, b, dx, dy = make_pair(cols, seed, m=rows)
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b))
    ref, _ = oracle.phase_correlate(a, b)
    assert (px, py) == (ref.peak_x, ref.peak_y) == (cols // 2 - dx, rows // 2 - dy)
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)
    assert resp == pytest.approx(ref.response, rel=1e-3, abs=1e-4)


def subpixel_shift(img, dx, dy):
    """Circular shift by a fractional (dx, dy): a phase ramp on the FP64
    spectrum (exact for a periodic band-limited image; the reference test
    uses cv::warpAffine with cubic interpolation, which needs OpenCV)."""
    m, n = img.shape
    ky = np.fft.fftfreq(m)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    ramp = np.exp(-2j * np.pi * (kx * dx + ky * dy))
    return np.fft.ifft2(np.fft.fft2(img.astype(np.float64)) * ramp).real.astype(np.float32)


def hanning_window(m, n):
    """Separable Hann window with the 2*pi*i/(N-1) phase of cv::createHanningWindow
    (the window UnitTest_glsPhaseCorrelate.cpp:96-97 applies)."""
    wy = 0.5 * (1.0 - np.cos(2 * np.pi * np.arange(m) / (m - 1)))
    wx = 0.5 * (1.0 - np.cos(2 * np.pi * np.arange(n) / (n - 1)))
    return np.outer(wy, wx).astype(np.float32)


@pytest.mark.parametrize("n,seed", [(32, 21), (64, 22), (256, 23), (512, 24), (1024, 25)])
def test_subpixel_shift_reference_bar(gls, n, seed):
    """UnitTest_glsPhaseCorrelate.cpp:45-147: a U[0,1) image, a sub-pixel
    shift within +-size/8, a Hann window; per axis |shift - gls| <= |shift -
    cv| + 0.25.  The oracle (the reference's chain restated in FP64) stands in
    for cv::phaseCorrelate; the device result must also match it to 1e-3 px."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 1, (n, n)).astype(np.float32)
    dx, dy = rng.uniform(-n / 8, n / 8, 2)
    b = subpixel_shift(a, dx, dy)
    w = hanning_window(n, n)
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b), gls.GlsMat(w))
    ref, _ = oracle.phase_correlate(a, b, w)
    print(f"n={n} truth=({dx:.3f},{dy:.3f}) gpu=({sx:.3f},{sy:.3f}) cpu=({ref.shift_x:.3f},{ref.shift_y:.3f})")
    assert (px, py) == (ref.peak_x, ref.peak_y)
    assert abs(dx - sx) <= abs(dx - ref.shift_x) + 0.25
    assert abs(dy - sy) <= abs(dy - ref.shift_y) + 0.25
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)
    assert resp == pytest.approx(ref.response, rel=1e-3, abs=1e-4)
    # the recovered shift is the true one to within a pixel (the integer peak is the nearest bin)
    assert abs(sx - dx) < 1.0 and abs(sy - dy) < 1.0


def test_exact_shift_is_a_delta(gls):
    rng = np.random.default_rng(10)
    a = rng.uniform(0, 1, (128, 128)).astype(np.float32)
    b = np.roll(a, (-17, 23), axis=(0, 1))
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b))
    assert (px, py) == (64 - 23, 64 + 17)
    assert sx == pytest.approx(23, abs=1e-3) and sy == pytest.approx(-17, abs=1e-3)
    assert resp == pytest.approx(1.0, abs=1e-3)


@pytest.mark.parametrize("scale", [1e-11, 1e-12])
def test_tiny_magnitudes_keep_the_peak(gls, scale):
    """The phase-only normalisation is scale-free: images scaled by 1e-11 or
    1e-12 (|A conj B|^2 in the denormal range for most bins) give the peak and
    shift of the unscaled pair, as the restated chain does."""
    a, b, dx, dy = make_pair(64, 12, noise=0.0)
    ref, _ = oracle.phase_correlate(a * scale, b * scale)
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat((a * scale).astype(np.float32)),
                                                  gls.GlsMat((b * scale).astype(np.float32)))
    assert (px, py) == (32 - dx, 32 - dy) == (ref.peak_x, ref.peak_y)
    assert np.isfinite([sx, sy, resp]).all()
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)


def test_zero_pair_gives_a_zero_surface(gls):
    """All-zero images: the cross power vanishes everywhere and is defined as
    0 (the reference divides 0/0); the chain reports what the restated chain
    reports for a zero surface."""
    z = np.zeros((64, 64), np.float32)
    ref, surf = oracle.phase_correlate(z, z, want_surface=True)
    assert not surf.any()
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(z), gls.GlsMat(z))
    assert (px, py) == (ref.peak_x, ref.peak_y)
    assert (sx, sy, resp) == pytest.approx((ref.shift_x, ref.shift_y, ref.response), abs=1e-9)


def test_peak_at_border_clamps_centroid(gls):
    # shift of -N/2 puts the peak on row/col 0 after SHIFT: the 5x5 box is clamped
    rng = np.random.default_rng(11)
    a = rng.uniform(0, 1, (64, 64)).astype(np.float32)
    b = np.roll(a, (32, 32), axis=(0, 1))
    (sx, sy), resp, (px, py) = gls.phaseCorrelate(gls.GlsMat(a), gls.GlsMat(b))
    ref, _ = oracle.phase_correlate(a, b)
    assert (px, py) == (ref.peak_x, ref.peak_y) == (0, 0)
    assert sx == pytest.approx(ref.shift_x, abs=1e-3) and sy == pytest.approx(ref.shift_y, abs=1e-3)


def test_bad_inputs_raise(gls):
    a = gls.GlsMat(np.zeros((64, 64), np.float32))
    with pytest.raises(gls.GlsError):
        gls.phaseCorrelate(a, gls.GlsMat(np.zeros((32, 64), np.float32)))
    with pytest.raises(gls.GlsError):
        gls.phaseCorrelate(a, gls.GlsMat(np.zeros((64, 64, 2), np.float32)))
    with pytest.raises(gls.GlsError):
        gls.phaseCorrelate(gls.GlsMat(np.zeros((48, 64), np.float32)), gls.GlsMat(np.zeros((48, 64), np.float32)))
