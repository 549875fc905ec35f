"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the oracle fails it:
brute-force O(N^2) DFTs on tiny inputs, numpy (pocketfft, FP64) golden
vectors committed under tests/golden/, and closed forms (delta, constant,
single tone, linearity, Parseval, round trip, flag identities).  A planted
mistake test shows the brute-force check catches an index error.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def brute_dft2(x, sign=-1):
    x = np.asarray(x, np.complex128)
    m, n = x.shape
    wm = np.exp(sign * 2j * np.pi * np.outer(np.arange(m), np.arange(m)) / m)
    wn = np.exp(sign * 2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    return wm @ x @ wn


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


SMALL = [(1, 1), (1, 2), (2, 1), (2, 2), (1, 8), (8, 1), (4, 4), (4, 8), (8, 16), (16, 16), (2, 32), (32, 4),
         (32, 32), (1, 64), (64, 1), (16, 64)]


@pytest.mark.parametrize("shape", SMALL)
def test_f64_matches_brute_force(shape):
    x = rnd(shape, 1)
    assert oracle.rel_rms(oracle.fft_f64(x), brute_dft2(x)) < 1e-13
    assert oracle.rel_rms(oracle.fft_f64(x, oracle.INVERSE), brute_dft2(x, +1)) < 1e-13


@pytest.mark.parametrize("shape", SMALL)
def test_r4_f32_matches_brute_force(shape):
    x = rnd(shape, 2)
    y = oracle.fft_r4_f32(x.astype(np.complex64))
    assert oracle.rel_rms(y, brute_dft2(x.astype(np.complex64))) < 2e-6


@pytest.mark.parametrize("name", ["16x16", "32x32", "1x32", "64x1", "64x128", "8x4", "cfg1_256x256"])
def test_golden_vectors(name):
    d = np.load(os.path.join(GOLD, f"fft_{name}.npz"))
    x = d["x"]
    fl = {"fwd": 0, "inv": oracle.INVERSE, "fwd_shift": oracle.SHIFT,
          "inv_scale": oracle.INVERSE | oracle.SCALE}
    for key, flags in fl.items():
        if key not in d:
            continue
        ref = d[key]
        assert oracle.rel_rms(oracle.fft_f64(x, flags), ref) < 1e-13, key
        assert oracle.rel_rms(oracle.fft_r4_f32(x, flags), ref) < 1e-6, key


def test_closed_forms():
    m, n = 16, 32
    d = np.zeros((m, n), np.complex128)
    d[0, 0] = 1
    assert np.allclose(oracle.fft_f64(d), 1.0)                       # delta -> ones
    c = np.full((m, n), 2.5 - 1j)
    y = oracle.fft_f64(c)
    assert np.isclose(y[0, 0], (2.5 - 1j) * m * n) and np.abs(y).ravel()[1:].max() < 1e-10
    k1, k2 = 3, 7                                                       # on-grid tone -> one bin
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    tone = np.exp(2j * np.pi * (k1 * ii / m + k2 * jj / n))
    y = oracle.fft_f64(tone)
    assert np.isclose(y[k1, k2], m * n) and np.sort(np.abs(y).ravel())[-2] < 1e-9
    a, b = rnd((m, n), 3), rnd((m, n), 4)                               # linearity
    assert np.allclose(oracle.fft_f64(2 * a - 3j * b), 2 * oracle.fft_f64(a) - 3j * oracle.fft_f64(b))
    ya = oracle.fft_f64(a)                                              # Parseval
    assert np.isclose(np.sum(np.abs(ya) ** 2), m * n * np.sum(np.abs(a) ** 2))


@pytest.mark.parametrize("impl", ["f64", "r4"])
@pytest.mark.parametrize("shape", [(16, 16), (8, 32), (1, 16), (32, 1), (64, 64)])
def test_flag_identities(impl, shape):
    x = rnd(shape, 5)
    f = (lambda v, fl=0: oracle.fft_f64(v, fl)) if impl == "f64" else \
        (lambda v, fl=0: oracle.fft_r4_f32(v.astype(np.complex64), fl).astype(np.complex128))
    tol = 1e-12 if impl == "f64" else 2e-6
    m, n = shape
    y = f(x)
    # INVERSE == conj o FFT o conj
    assert oracle.rel_rms(f(x, oracle.INVERSE), np.conj(f(np.conj(x)))) < tol
    # SCALE == x 1/(M N)
    assert oracle.rel_rms(f(x, oracle.SCALE), y / (m * n)) < tol
    # SHIFT == quadrant swap (the test's fftShift, unittest_glsFft.cpp:44-97)
    assert oracle.rel_rms(f(x, oracle.SHIFT), np.roll(y, (m // 2, n // 2), axis=(0, 1))) < tol
    # forward -> inverse(SCALE) round trip
    assert oracle.rel_rms(f(y, oracle.INVERSE | oracle.SCALE), x) < tol


def test_r4_restatement_odd_and_even_stage_counts():
    # N in {64, 128, 1024, 2048} have an odd number of radix-4 stages, the case
    # where the reference's shared `bank` toggling went wrong.
    for n in (64, 128, 512, 1024, 2048):
        x = rnd((4, n), n)
        assert oracle.rel_rms(oracle.fft_r4_f32(x), np.fft.fft2(x)) < 1e-6, n


@pytest.mark.parametrize("fault", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(fault):
    """Plant a known mistake in the oracle and check that the pinning checks
    above reject it (a wrong twiddle index, a wrong sign in one dimension, a
    half-applied SHIFT, a dropped conjugate in the POC)."""
    x = rnd((8, 16), 6)
    oracle.plant_fault(fault)
    try:
        if fault == 1:
            bad = oracle.rel_rms(oracle.fft_r4_f32(x.astype(np.complex64)), brute_dft2(x))
        elif fault == 2:
            bad = oracle.rel_rms(oracle.fft_f64(x), brute_dft2(x))
        elif fault == 3:
            bad = oracle.rel_rms(oracle.fft_f64(x, oracle.SHIFT), np.roll(brute_dft2(x), (4, 8), axis=(0, 1)))
        else:
            a, b = rnd((8, 8), 7), rnd((8, 8), 8)
            bad = float(np.max(np.abs(np.angle(oracle.poc(a, b) / (a * np.conj(b))))))
    finally:
        oracle.plant_fault(0)
    assert bad > 1e-3
    # and with the fault removed the same check passes
    assert oracle.rel_rms(oracle.fft_f64(x), brute_dft2(x)) < 1e-13


def test_poc_closed_form():
    a, b = rnd((8, 8), 7), rnd((8, 8), 8)
    c = oracle.poc(a, b)
    assert np.allclose(np.abs(c), 1.0, atol=1e-6)
    assert np.allclose(np.angle(c), np.angle(a * np.conj(b)), atol=1e-5)
    z = oracle.poc(np.zeros((2, 2)), np.zeros((2, 2)))                 # epsilon guard: 0/0 -> 0
    assert np.all(z == 0)


def test_argmax_first_maximum_row_major():
    v = np.zeros((4, 5), np.float32)
    v[2, 1] = v[1, 3] = v[3, 0] = 7.0
    assert oracle.argmax_first(v)[:2] == (3, 1)                         # (x, y) of row 1 wins
    v[0, 4] = np.nan
    assert oracle.argmax_first(v)[:2] == (3, 1)                         # NaN never wins


def test_weighted_centroid_closed_form():
    v = np.zeros((16, 16), np.float32)
    v[5, 7] = 3.0
    v[5, 8] = 1.0
    cx, cy, r = oracle.weighted_centroid(v, 7, 5)
    assert np.isclose(cx, (7 * 3 + 8 * 1) / 4.0) and np.isclose(cy, 5.0) and np.isclose(r, 4.0)
    v2 = np.zeros((8, 8), np.float32)                                   # clamped at the border
    v2[0, 0] = 2.0
    v2[7, 7] = 100.0                                                    # outside the clamped box
    cx, cy, r = oracle.weighted_centroid(v2, 0, 0)
    assert (cx, cy, r) == pytest.approx((0.0, 0.0, 2.0))


@pytest.mark.parametrize("n,dx,dy", [(32, 3, -5), (64, -7, 12), (128, 20, -31), (64, 0, 0)])
def test_phase_correlate_integer_shift(n, dx, dy):
    rng = np.random.default_rng(n + dx)
    img = rng.uniform(0, 1, (n, n)).astype(np.float32)
    img2 = np.roll(img, (dy, dx), axis=(0, 1))
    res, surf = oracle.phase_correlate(img, img2, want_surface=True)
    # exact circular shift: surface = M N delta at (N/2 - d) after SHIFT
    assert (res.peak_x, res.peak_y) == (n // 2 - dx, n // 2 - dy)
    assert res.shift_x == pytest.approx(dx, abs=1e-3) and res.shift_y == pytest.approx(dy, abs=1e-3)
    assert res.response == pytest.approx(1.0, abs=1e-3)
    assert surf[res.peak_y, res.peak_x] == pytest.approx(n * n, rel=1e-4)


# ---- spectrum display chain (glsBasicOperation.cpp:337-351, glsNormalize.cpp:52-75) ----

def test_mag_and_log_mag_closed_forms():
    c = np.array([3 + 4j, -5 - 12j, 0, 1e-3j, -2 + 0j], np.complex64)
    assert np.allclose(oracle.mag_spectrums(c), [5, 13, 0, 1e-3, 2], rtol=1e-7, atol=0)
    lm = oracle.log_mag_spectrums(c, 1.0)
    assert np.allclose(lm, np.log([6, 14, 1, 1.001, 3]), rtol=1e-6, atol=1e-7)
    # |.| inside the log: a negative offset larger than the magnitude
    assert np.allclose(oracle.log_mag_spectrums(np.array([1 + 0j], np.complex64), -3.0), [np.log(2.0)], rtol=1e-6)
    # mag of a random plane against numpy in float64
    rng = np.random.default_rng(3)
    z = (rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))).astype(np.complex64)
    assert np.allclose(oracle.mag_spectrums(z), np.abs(z.astype(np.complex128)), rtol=2e-7, atol=0)


def test_normalize_minmax_maps_extremes():
    rng = np.random.default_rng(4)
    x = rng.uniform(-3, 7, (32, 16)).astype(np.float32)
    y = oracle.normalize(x, 0.0, 1.0, oracle.NORM_MINMAX)
    i_min, i_max = np.argmin(x), np.argmax(x)
    assert abs(y.flat[i_min] - 0.0) < 1e-6 and abs(y.flat[i_max] - 1.0) < 1e-6
    assert np.all(y >= -1e-6) and np.all(y <= 1 + 1e-6)
    # affine: same ordering, same ratios as float64
    ref = (x.astype(np.float64) - x.min()) / (float(x.max()) - float(x.min()))
    assert np.allclose(y, ref, atol=3e-7)
    # alpha > beta reverses, as the reference's formula does
    r = oracle.normalize(x, 2.0, -1.0, oracle.NORM_MINMAX)
    assert abs(r.flat[i_min] - 2.0) < 1e-5 and abs(r.flat[i_max] + 1.0) < 1e-5
    # constant plane: alpha everywhere (no division by zero)
    assert np.all(oracle.normalize(np.full((4, 4), 3.0, np.float32), 0.5, 1.0, oracle.NORM_MINMAX) == 0.5)


def test_normalize_l1_l2_norms():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((8, 64)).astype(np.float32)
    y2 = oracle.normalize(x, 2.0, 0.0, oracle.NORM_L2)
    assert abs(np.linalg.norm(y2.astype(np.float64)) - 2.0) < 1e-5
    y1 = oracle.normalize(x, 3.0, 0.0, oracle.NORM_L1)
    assert abs(np.abs(y1.astype(np.float64)).sum() - 3.0) < 1e-5
    with pytest.raises(ValueError):
        oracle.normalize(x, 1.0, 0.0, 1)   # NORM_INF is not a normalize type in the reference


@pytest.mark.parametrize("fault", [5, 6])
def test_planted_display_chain_mistakes_are_caught(fault):
    c = np.array([3 + 4j, 0.5 + 0j, 1j], np.complex64)
    x = np.array([1.0, 5.0, 3.0], np.float32)
    oracle.plant_fault(fault)
    try:
        if fault == 5:
            ok = np.allclose(oracle.log_mag_spectrums(c, 1.0), np.log([6, 1.5, 2]), rtol=1e-6)
        else:
            y = oracle.normalize(x, 0.0, 1.0, oracle.NORM_MINMAX)
            ok = abs(y[0]) < 1e-6 and abs(y[1] - 1) < 1e-6
    finally:
        oracle.plant_fault(0)
    assert not ok


# ---- the chain's standalone operations (mulSpectrums, minMaxLoc) ----------------
FLT_MAX = np.float32(np.finfo(np.float32).max)


def minmax_brute(v):
    """gls::minMaxLoc's result by another route: numpy's min / max over the
    values the strict scans can pick (not NaN, past the start values
    FLT_MAX / -FLT_MAX), then the first index holding it in row-major order;
    (start value, (0, 0)) when no value qualifies."""
    v = np.asarray(v, np.float32)
    ch0 = v[..., 0] if v.ndim == 3 else v
    cols = ch0.shape[1]
    flat = ch0.reshape(-1)
    out = []
    for pick, ok, start in ((np.min, ~np.isnan(flat) & (flat < FLT_MAX), FLT_MAX),
                            (np.max, ~np.isnan(flat) & (flat > -FLT_MAX), -FLT_MAX)):
        if ok.any():
            m = pick(flat[ok])
            i = int(np.flatnonzero(ok & (flat == m))[0])
            out.append((float(m), (i % cols, i // cols)))
        else:
            out.append((float(start), (0, 0)))
    return out[0][0], out[1][0], out[0][1], out[1][1]


def minmax_cases():
    rng = np.random.default_rng(41)
    cases = []
    for shape in ((1, 1), (1, 9), (9, 1), (3, 5), (16, 16), (37, 53)):
        cases.append(rng.standard_normal(shape).astype(np.float32))
        cases.append(rng.integers(-3, 4, shape).astype(np.float32))          # many ties
    x = rng.integers(-2, 3, (12, 10)).astype(np.float32)
    x[rng.random(x.shape) < 0.3] = np.nan
    cases.append(x)
    y = rng.standard_normal((8, 8)).astype(np.float32)
    y[2, 3], y[5, 1], y[0, 7] = np.inf, -np.inf, np.nan
    cases.append(y)
    cases.append(np.full((4, 6), np.nan, np.float32))
    cases.append(np.full((4, 6), -np.inf, np.float32))                        # max: -FLT_MAX at (0, 0)
    cases.append(np.full((4, 6), np.inf, np.float32))                         # min: FLT_MAX at (0, 0)
    z = np.zeros((5, 7), np.float32)
    z[3, 2], z[1, 6] = -FLT_MAX, FLT_MAX                                       # the start values themselves
    cases.append(z)
    cases.append(np.full((6, 6), -FLT_MAX, np.float32))
    c2 = rng.standard_normal((9, 11, 2)).astype(np.float32)
    c2[..., 1] *= 100.0                                                        # channel 1 is ignored
    cases.append(c2)
    return cases


def test_min_max_loc_matches_brute_force():
    for k, v in enumerate(minmax_cases()):
        got = oracle.min_max_loc(v)
        want = minmax_brute(v)
        assert got == want, (k, got, want)


def test_min_max_loc_agrees_with_argmax_first():
    rng = np.random.default_rng(5)
    for shape in ((7, 9), (32, 32), (1, 100)):
        v = rng.integers(0, 5, shape).astype(np.float32)
        x, y, m = oracle.argmax_first(v)
        assert oracle.min_max_loc(v)[1:4:2] == (m, (x, y))


def test_mul_spectrums_closed_forms():
    a = rnd((6, 8), 11).astype(np.complex64)
    b = rnd((6, 8), 12).astype(np.complex64)
    one = np.ones_like(a)
    assert np.array_equal(oracle.mul_spectrums(a, one).view(np.uint32), a.view(np.uint32))
    i = np.full_like(a, 1j)
    assert np.array_equal(oracle.mul_spectrums(a, i), -a.imag + 1j * a.real)      # a*i = (-ai, ar) exactly
    aa = oracle.mul_spectrums(a, a, conj=True)
    assert np.all(aa.imag == 0)                                                     # a*conj(a) is real
    # a*conj(b) == a*(conj b), bit for bit
    assert np.array_equal(oracle.mul_spectrums(a, b, conj=True).view(np.uint32),
                          oracle.mul_spectrums(a, np.conj(b)).view(np.uint32))
    for conj in (False, True):
        ref = a.astype(np.complex128) * (np.conj(b) if conj else b).astype(np.complex128)
        assert oracle.rel_rms(oracle.mul_spectrums(a, b, conj), ref) < 1e-6


@pytest.mark.parametrize("fault", [7, 8])
def test_planted_chain_op_mistakes_are_caught(fault):
    a = rnd((4, 4), 3).astype(np.complex64)
    v = np.array([[1, 9, 9], [0, 9, 2]], np.float32)
    oracle.plant_fault(fault)
    try:
        if fault == 7:
            ok = np.array_equal(oracle.mul_spectrums(a, np.full_like(a, 1j)), -a.imag + 1j * a.real)
        else:
            ok = oracle.min_max_loc(v) == minmax_brute(v)
    finally:
        oracle.plant_fault(0)
    assert not ok
