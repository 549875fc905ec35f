"""GPU tests of the batched plans (one persistent task-graph launch for many
FFTs / phase correlations of different shapes).

Each batched result must be bit-identical to the same transform run alone
through the plain path (same per-transform arithmetic; only the tiling of
the launch differs) and within 1e-5 rel-RMS of an FP64 FFT.  Plans transform
a real (CV_32FC1) entry as the complex image with imaginary part 0, so its
bits are those of the plain path's complex transform of that image (the
plain path's real-input transform runs half-size spectra instead).  Graph replays
are checked one by one against the eager result, with other device work in
between."""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
SIZES = [64, 128, 256, 512, 1024, 2048]


def rnd_c(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


def make_batch(gls, shapes, flags_list=None, seed=0, inplace=False, real_idx=()):
    srcs, dsts, descs, xs = [], [], [], []
    for i, (r, c) in enumerate(shapes):
        fl = flags_list[i] if flags_list else 0
        if i in real_idx:
            x = np.random.default_rng(seed + i).uniform(-1, 1, (r, c)).astype(np.float32)
            src = gls.GlsMat(x)
            st = gls.CV_32FC1
        else:
            x = rnd_c((r, c), seed + i)
            src = gls.GlsMat(x)
            st = gls.CV_32FC2
        dst = src if inplace and st == gls.CV_32FC2 else gls.GlsMat((r, c), gls.CV_32FC2)
        srcs.append(src)
        dsts.append(dst)
        xs.append(x)
        descs.append(gls.FftDesc(src.dev_ptr, dst.dev_ptr, r, c, st, fl))
    return srcs, dsts, descs, xs


def plain(gls, x, flags):
    return gls.fft(gls.GlsMat(x), None, flags).download_complex()


@pytest.mark.parametrize("mode", [0, 2])
def test_uneven_batch_bit_identical_to_plain_and_fp64(gls, mode):
    rng = np.random.default_rng(3)
    shapes = [(int(rng.choice(SIZES)), int(rng.choice(SIZES))) for _ in range(24)]
    srcs, dsts, descs, xs = make_batch(gls, shapes, seed=100)
    plan = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=256, persistent=mode)
    plan.execute()
    plan.check()
    for (r, c), x, d in zip(shapes, xs, dsts):
        y = d.download_complex()
        assert np.array_equal(y.view(np.uint32), plain(gls, x, 0).view(np.uint32)), (r, c)
        assert oracle.rel_rms(y, np.fft.fft2(x.astype(np.complex128))) <= 1e-5, (r, c)


@pytest.mark.parametrize("mode", [0, 2])
def test_mixed_flags_1d_real_and_inplace(gls, mode):
    shapes = [(1, 4096), (2048, 1), (64, 2048), (512, 64), (256, 256), (128, 1024), (1, 2), (4, 4)]
    flags = [0, 1, 3, 64, 67, 65, 2, 66]
    srcs, dsts, descs, xs = make_batch(gls, shapes, flags, seed=7, real_idx=(3,))
    plan = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=256, persistent=mode)
    plan.execute()
    plan.check()
    for (r, c), fl, x, d in zip(shapes, flags, xs, dsts):
        y = d.download_complex()
        xin = x.astype(np.complex64) if x.dtype == np.float32 else x
        assert np.array_equal(y.view(np.uint32), plain(gls, xin, fl).view(np.uint32)), (r, c, fl)
        assert oracle.rel_rms(y, oracle.fft_f64(x.astype(np.complex128), fl)) <= 1e-5, (r, c, fl)
    # in place: every entry transforms its own buffer
    shapes2 = [(128, 512), (1024, 64), (64, 64)]
    srcs2, dsts2, descs2, xs2 = make_batch(gls, shapes2, seed=9, inplace=True)
    p2 = gls.Plan.fft_batch(descs2, 0, keep=(srcs2, dsts2), threads=256, persistent=mode)
    p2.execute()
    p2.check()
    for x, d in zip(xs2, dsts2):
        assert oracle.rel_rms(d.download_complex(), np.fft.fft2(x.astype(np.complex128))) <= 1e-5


@pytest.mark.parametrize("mode", [0, 2])
def test_wave_exchange_rows_bit_identical_to_lds_plans(gls, mode):
    """Row lengths 32, 64, 512 and 1024 exchange inside the wave on the plain
    path (DPP, ds_swizzle, permlane swaps); the plan kernels exchange through
    LDS.  Same bits for every flag, real input, 1D rows and partial waves
    (2 or 4 rows of 32 points: a 4- or 8-lane workgroup)."""
    shapes = [(1, 32), (2, 32), (4, 32), (64, 32), (1, 64), (2, 64), (256, 64), (1, 512), (2, 512),
              (64, 512), (1, 1024), (4, 1024), (128, 1024), (1024, 1024)]
    flags = [0, 1, 3, 64, 67, 2, 65, 1, 64, 3, 66, 0, 67, 0]
    srcs, dsts, descs, xs = make_batch(gls, shapes, flags, seed=31, real_idx=(3, 9, 12))
    plan = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=256, persistent=mode)
    plan.execute()
    plan.check()
    for (r, c), fl, x, d in zip(shapes, flags, xs, dsts):
        y = d.download_complex()
        xin = x.astype(np.complex64) if x.dtype == np.float32 else x
        assert np.array_equal(y.view(np.uint32), plain(gls, xin, fl).view(np.uint32)), (r, c, fl)
        assert oracle.rel_rms(y, oracle.fft_f64(x.astype(np.complex128), fl)) <= 1e-5, (r, c, fl)


def test_graph_replays_match_eager(gls):
    rng = np.random.default_rng(5)
    shapes = [(int(rng.choice(SIZES[:5])), int(rng.choice(SIZES[:5]))) for _ in range(16)]
    srcs, dsts, descs, xs = make_batch(gls, shapes, seed=200)
    eager = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts))
    eager.execute()
    eager.check()
    ref = [d.download_complex().copy() for d in dsts]
    graph = gls.Plan.fft_batch(descs, gls.PLAN_GRAPH, keep=(srcs, dsts))
    assert graph.info().uses_graph == 1
    pgraph = gls.Plan.fft_batch(descs, gls.PLAN_GRAPH, keep=(srcs, dsts), threads=256, persistent=1)
    assert pgraph.info().uses_graph == 1 and pgraph.info().persistent == 1
    for rep in range(3):                                # persistent kernel: counters reset per replay
        for d in dsts:
            d.upload(np.zeros((d.rows, d.cols), np.complex64))
        pgraph.execute()
        pgraph.check()
        for r_, d in zip(ref, dsts):
            assert np.array_equal(d.download_complex().view(np.uint32), r_.view(np.uint32)), rep
    for g in (graph, gls.Plan.fft_batch(descs, gls.PLAN_GRAPH, keep=(srcs, dsts), threads=256, persistent=2)):
        for rep in range(4):
            for d in dsts:                              # clobber the outputs with other device work
                gls.fft(gls.GlsMat(np.ones((8, 8), np.complex64)), None, 0)
                d.upload(np.zeros((d.rows, d.cols), np.complex64))
            g.execute()
            g.check()
            for r_, d in zip(ref, dsts):
                assert np.array_equal(d.download_complex().view(np.uint32), r_.view(np.uint32)), rep


@pytest.mark.parametrize("persistent", [0, 1, 2])
@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_launch_shapes_agree(gls, threads, persistent):
    """Every workgroup size, in every execution mode, gives the same bits
    (grouped levels take each pass's shape from the plain path)."""
    shapes = [(2048, 64), (64, 2048), (512, 512), (4096, 16), (16, 4096)]
    srcs, dsts, descs, xs = make_batch(gls, shapes, seed=300)
    plan = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=threads, col_wcap=4 if threads == 512 else 16,
                              persistent=persistent)
    assert plan.info().threads == threads and plan.info().persistent == persistent
    plan.execute()
    plan.check()
    for x, d in zip(xs, dsts):
        assert np.array_equal(d.download_complex().view(np.uint32), plain(gls, x, 0).view(np.uint32))


def test_autotune_and_database(gls, tmp_path, monkeypatch):
    db = tmp_path / "tune.db"
    monkeypatch.setenv("GLSX_TUNE_DB", str(db))
    rng = np.random.default_rng(11)
    shapes = [(int(rng.choice(SIZES)), int(rng.choice(SIZES))) for _ in range(12)]
    srcs, dsts, descs, xs = make_batch(gls, shapes, seed=400, inplace=True)
    plan = gls.Plan.fft_batch(descs, gls.PLAN_AUTOTUNE, keep=(srcs, dsts))
    info = plan.info()
    assert info.tuned == 1 and info.tuned_us > 0
    assert db.exists() and len(db.read_text().strip().splitlines()) == 1
    # tuning ran on scratch buffers: the in-place inputs are untouched
    for x, s in zip(xs, srcs):
        assert np.array_equal(s.download_complex(), x)
    plan.execute()
    plan.check()
    for x, d in zip(xs, dsts):
        assert oracle.rel_rms(d.download_complex(), np.fft.fft2(x.astype(np.complex128))) <= 1e-5
    again = gls.Plan.fft_batch(descs, gls.PLAN_AUTOTUNE, keep=(srcs, dsts))
    assert again.info().tuned == 2                       # served from the database
    assert (again.info().threads, again.info().col_wcap) == (info.threads, info.col_wcap)
    # a damaged entry for the same key (a column width no tile can use) is
    # skipped: the earlier valid entry is still served
    key = db.read_text().split()[0]
    with open(db, "a") as f:
        f.write(f"{key} 512 3 0 1.0\n{key} 333 16 0 1.0\n{key} 512 16 7 1.0\n")
    third = gls.Plan.fft_batch(descs, gls.PLAN_AUTOTUNE, keep=(srcs, dsts))
    assert third.info().tuned == 2
    assert (third.info().threads, third.info().col_wcap) == (info.threads, info.col_wcap)


def test_explicit_config_is_validated(gls):
    srcs, dsts, descs, xs = make_batch(gls, [(64, 64), (128, 32)], seed=17)
    for threads, wcap, pers in ((512, 3, 0), (512, 0, 0), (512, 128, 0), (512, 16, 3), (512, 16, -1), (300, 16, 0)):
        with pytest.raises(gls.GlsError):
            gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=threads, col_wcap=wcap, persistent=pers)
    ok = gls.Plan.fft_batch(descs, 0, keep=(srcs, dsts), threads=512, col_wcap=8, persistent=0)
    ok.execute()
    ok.check()
    for x, d in zip(xs, dsts):
        assert np.array_equal(d.download_complex().view(np.uint32), plain(gls, x, 0).view(np.uint32))


def box5(img):
    acc = np.zeros_like(img, dtype=np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            acc += np.roll(img, (dy, dx), axis=(0, 1))
    return (acc / 25.0).astype(np.float32)


def poc_pair(n, seed):
    rng = np.random.default_rng(seed)
    a = box5(rng.integers(0, 256, (n, n), dtype=np.uint8).astype(np.float32) / 255.0)
    dx, dy = int(rng.integers(-n // 4, n // 4)), int(rng.integers(-n // 4, n // 4))
    b = (np.roll(a, (dy, dx), axis=(0, 1)) + rng.normal(0, 0.05, (n, n))).astype(np.float32)
    return a, b, dx, dy


@pytest.mark.parametrize("threads,persistent", [(256, 0), (1024, 0), (256, 1), (512, 1), (256, 2)])
def test_poc_batch_same_alone_and_in_the_mix(gls, threads, persistent):
    sizes = [256, 2048, 512, 1024, 256, 512, 1024, 256]
    pairs = [poc_pair(n, 50 + i) for i, n in enumerate(sizes)]
    mats = [(gls.GlsMat(a), gls.GlsMat(b)) for a, b, _, _ in pairs]
    descs = [gls.PocDesc(ma.dev_ptr, mb.dev_ptr, None, n, n) for (ma, mb), n in zip(mats, sizes)]
    ws_bytes = gls.Plan.poc_workspace_bytes(descs)
    ws = gls.GlsMat((1, ws_bytes), gls.CV_8UC1)
    plan = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, threads, 0, keep=(mats, ws), persistent=persistent)
    plan.execute()
    plan.check()
    res = plan.results()
    for (a, b, dx, dy), n, r, (ma, mb) in zip(pairs, sizes, res, mats):
        alone = gls.phase_correlate_dev(ma.dev_ptr, mb.dev_ptr, None, n, n)
        assert (r.peak_x, r.peak_y) == (n // 2 - dx, n // 2 - dy)
        assert (r.peak_x, r.peak_y, r.shift_x, r.shift_y, r.response) == \
               (alone.peak_x, alone.peak_y, alone.shift_x, alone.shift_y, alone.response)
    # and again, reusing the plan (counters reset by the kernel itself)
    plan.execute()
    plan.check()
    res2 = plan.results()
    assert [(r.peak_x, r.peak_y, r.shift_x) for r in res2] == [(r.peak_x, r.peak_y, r.shift_x) for r in res]


def test_poc_autotune_matches_default_shape(gls, tmp_path, monkeypatch):
    db = tmp_path / "tune.db"
    monkeypatch.setenv("GLSX_TUNE_DB", str(db))
    sizes = [256, 512, 1024, 256, 2048, 512]
    pairs = [poc_pair(n, 70 + i) for i, n in enumerate(sizes)]
    mats = [(gls.GlsMat(a), gls.GlsMat(b)) for a, b, _, _ in pairs]
    descs = [gls.PocDesc(ma.dev_ptr, mb.dev_ptr, None, n, n) for (ma, mb), n in zip(mats, sizes)]
    ws_bytes = gls.Plan.poc_workspace_bytes(descs)
    ws = gls.GlsMat((1, ws_bytes), gls.CV_8UC1)
    tuned = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, gls.PLAN_AUTOTUNE, keep=(mats, ws))
    info = tuned.info()
    assert info.tuned == 1 and info.tuned_us > 0
    assert db.exists() and "pocbatch|" in db.read_text()
    tuned.execute()
    tuned.check()
    got = tuned.results()
    plain = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, 0, keep=(mats, ws))
    assert plain.info().threads == 512
    plain.execute()
    plain.check()
    ref = plain.results()
    # every launch shape computes the same bits
    for r, q, (a, b, dx, dy), n in zip(got, ref, pairs, sizes):
        assert (r.peak_x, r.peak_y) == (n // 2 - dx, n // 2 - dy)
        assert (r.peak_x, r.peak_y, r.shift_x, r.shift_y, r.response) == \
               (q.peak_x, q.peak_y, q.shift_x, q.shift_y, q.response)
    again = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, gls.PLAN_AUTOTUNE, keep=(mats, ws))
    assert again.info().tuned == 2


def test_poc_workgroup_too_small_rejected(gls):
    # row-pair tiles of 4096-wide images need 512 threads
    a = gls.GlsMat(np.zeros((16, 4096), np.float32))
    descs = [gls.PocDesc(a.dev_ptr, a.dev_ptr, None, 16, 4096)]
    ws_bytes = gls.Plan.poc_workspace_bytes(descs)
    ws = gls.GlsMat((1, ws_bytes), gls.CV_8UC1)
    with pytest.raises(gls.GlsError):
        gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 256, 0, keep=(a, ws))
    ok = gls.Plan.poc_batch(descs, ws.dev_ptr, ws_bytes, 0, 0, keep=(a, ws))
    assert ok.info().threads == 512


def test_bad_plans_rejected(gls):
    m = gls.GlsMat((64, 8192), gls.CV_32FC2)
    with pytest.raises(gls.GlsError):
        gls.Plan.fft_batch([gls.FftDesc(m.dev_ptr, m.dev_ptr, 64, 8192, gls.CV_32FC2, 0)])
    with pytest.raises(gls.GlsError):
        gls.Plan.fft_batch([gls.FftDesc(m.dev_ptr, m.dev_ptr, 48, 64, gls.CV_32FC2, 0)])
    with pytest.raises(gls.GlsError):
        gls.Plan.fft_batch([gls.FftDesc(None, m.dev_ptr, 64, 64, gls.CV_32FC2, 0)])


def test_config3_full_batch_every_transform_vs_fp64(gls):
    """The benchmark's batched workload exactly (256 transforms, rows and cols
    drawn from {64..2048} with the bench's seed, entries U[-1,1]), as the
    bench runs it (autotuned launch shape, captured graph), with every one of
    the 256 results checked against numpy's FP64 FFT."""
    rng = np.random.default_rng(20261203)
    shapes = [(int(rng.choice(SIZES)), int(rng.choice(SIZES))) for _ in range(256)]
    srcs, dsts, descs, xs = [], [], [], []
    for r, c in shapes:
        x = (rng.uniform(-1, 1, (r, c)) + 1j * rng.uniform(-1, 1, (r, c))).astype(np.complex64)
        a, b = gls.GlsMat(x), gls.GlsMat((r, c), gls.CV_32FC2)
        srcs.append(a)
        dsts.append(b)
        xs.append(x)
        descs.append(gls.FftDesc(a.dev_ptr, b.dev_ptr, r, c, gls.CV_32FC2, 0))
    plan = gls.Plan.fft_batch(descs, gls.PLAN_GRAPH | gls.PLAN_AUTOTUNE, keep=(srcs, dsts))
    plan.execute()
    plan.check()
    worst = 0.0
    for (r, c), x, d in zip(shapes, xs, dsts):
        ref = np.fft.fft2(x.astype(np.complex128))
        err = oracle.rel_rms(d.download_complex(), ref)
        assert err <= 1e-5, (r, c, err)
        worst = max(worst, err)
    print(f"256 transforms, worst rel-RMS {worst:.3e}")
    plan.destroy()
