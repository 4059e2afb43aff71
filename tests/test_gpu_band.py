"""Band search around a caller's initial disparity map (MatcherUseDisparity) and the
disparity mip maps (CreateDispMipMaps) against the CPU oracle."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu


def _tile_bins(init, W, H, D, ref_format=False):
    """Per-pixel bins of the band search: every 64x56 tile keeps the 8-disparity bins of
    [min - 4, max + 4] over scale * d at (x / scale, y / scale), all bins when none is valid."""
    ih, iw = init.shape[:2]
    s = W // iw
    if ref_format:
        ids = init[..., 0].astype(np.int64)
        d = np.where((ids >= 1) & (ids <= 0xFFEF), ids - 1, -1)
    else:
        d = init.astype(np.int64)
    ys = np.minimum(np.arange(H) // s, ih - 1)
    xs = np.minimum(np.arange(W) // s, iw - 1)
    full = d[ys][:, xs]
    allm = (1 << (D // 8)) - 1
    bins = np.zeros((H, W), np.uint32)
    for y0 in range(0, H, 56):
        for x0 in range(0, W, 64):
            blk = full[y0:y0 + 56, x0:x0 + 64]
            v = blk[blk >= 0] * s
            if v.size == 0:
                m = allm
            else:
                lo, hi = max(0, int(v.min()) - 4), min(D - 1, int(v.max()) + 4)
                m = 0
                for b in range(lo >> 3, (hi >> 3) + 1):
                    m |= 1 << b
                m &= allm
            bins[y0:y0 + 56, x0:x0 + 64] = m
    return bins


def _oracle(algo, L, R, D, bins, window, census):
    if algo == cs.ALGO_CENSUS:
        return orc.bm_bins(orc.COST_CENSUS, orc.census(L, *census), orc.census(R, *census), D, bins, window=window,
                           census_win=census, tol=1, nthreads=8)
    return orc.bm_bins(orc.COST_SAD, L, R, D, bins, window=window, tol=1, nthreads=8)


@pytest.mark.parametrize("algo,window,scale,fmt", [(cs.ALGO_CENSUS, (9, 9), 2, "i16"), (cs.ALGO_CENSUS, (9, 9), 1, "ref"),
                                                   (cs.ALGO_SAD, (7, 7), 3, "i16"), (cs.ALGO_SAD, (9, 9), 2, "ref")])
def test_band_search_matches_oracle(ctx, algo, window, scale, fmt):
    W, H, D = 400, 250, 128
    print(f"band search algo {algo} window {window} scale {scale} init {fmt}", flush=True)
    L, R = cs.synth_pair(91, W, H, 2, 8, D - 1, 8)
    census = (9, 7)
    full = cs.Matcher(ctx, W, H, algorithm=algo, window=window, census=census, num_disp=D, tol=1)
    ref_full = full.run(L, R)
    # the initial map: the full-range result at 1/scale (every scale-th pixel, d / scale),
    # with some values made invalid
    rng = np.random.default_rng(scale)
    small = ref_full[::scale, ::scale][: H // scale, : W // scale] // scale
    small = np.where(rng.random(small.shape) < 0.1, -1, small).astype(np.int16)
    if fmt == "ref":
        init = np.zeros(small.shape + (4,), np.uint16)
        init[..., 0] = np.where(small >= 0, small + 1, 0xFFFA)
    else:
        init = small
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=window, census=census, num_disp=D, tol=1, use_disp_input=True)
    got = m.run_init(L, R, init, maps=True)
    bins = _tile_bins(init, W, H, D, ref_format=(fmt == "ref"))
    want = _oracle(algo, L, R, D, bins, window, census)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    # most of the frame keeps the full-range result
    print(f"  agreement with the full range {(got[0] == ref_full).mean():.4f}", flush=True)


@pytest.mark.parametrize("scale,win,fmt,size", [(2, (11, 11), "i16", (400, 250, 128)), (1, (7, 7), "ref", (400, 250, 128)),
                                               (4, (9, 9), "i16", (400, 250, 128)), (3, (5, 5), "i16", (333, 190, 64))])
def test_ncc_band_search_matches_oracle(ctx, scale, win, fmt, size):
    """NCC around an initial map (MatcherUseDisparity of BlockMatcherNCC): the oracle's NCC
    restricted to the same per-pixel bins, with the NCC matcher's parity rules: the
    GPU's winning scores within 1e-5 relative of FP64, maps equal except at FP32
    near-ties, and every winner inside its pixel's bins."""
    W, H, D = size
    print(f"ncc band search scale {scale} window {win} init {fmt} size {size}", flush=True)
    L, R = cs.synth_pair(95, W, H, 2, 8, D - 1, 8)
    full = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_NCC, window=win, num_disp=D, tol=1)
    ref_full = full.run(L, R)
    rng = np.random.default_rng(10 + scale)
    small = ref_full[::scale, ::scale][: H // scale, : W // scale] // scale
    small = np.where(rng.random(small.shape) < 0.1, -1, small).astype(np.int16)
    small[: 60 // scale, : 130 // scale] = -1  # tiles with no valid value search the full range
    if fmt == "ref":
        init = np.zeros(small.shape + (4,), np.uint16)
        init[..., 0] = np.where(small >= 0, small + 1, 0xFFFD)
    else:
        init = small
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_NCC, window=win, num_disp=D, tol=1, use_disp_input=True)
    out, lr, rl, score = m.run_init(L, R, init, maps=True, scores=True)
    bins = _tile_bins(init, W, H, D, ref_format=(fmt == "ref"))
    ref_out, ref_l, ref_r, _ = orc.ncc_bins(L, R, D, bins, window=win, tol=1, nthreads=8)
    vol = orc.ncc(L, R, D, window=win, tol=1, nthreads=8, with_volume=True)[4]
    valid = ref_l >= 0
    assert np.array_equal(valid, lr >= 0)
    yy, xx = np.nonzero(valid)
    assert ((bins[yy, xx] >> (lr[yy, xx] >> 3).astype(np.uint32)) & 1).all()
    s64 = vol[yy, xx, lr[yy, xx]]
    s32 = score[yy, xx].astype(np.float64)
    zero = s64 == 0.0
    assert (s32[zero] == 0.0).all()
    assert (np.abs(s32[~zero] - s64[~zero]) / np.abs(s64[~zero])).max() <= 1e-5
    for a, b in ((lr, ref_l), (rl, ref_r), (out, ref_out)):
        assert (a == b).mean() >= 0.999
    dif = np.nonzero((lr != ref_l) & valid)
    if dif[0].size:
        assert np.all(np.abs(vol[dif[0], dif[1], lr[dif]] - vol[dif[0], dif[1], ref_l[dif]]) <= 1e-6)
    searched = np.mean([bin(int(b)).count("1") for b in np.unique(bins)]) * 8 / D
    print(f"  agreement with the full range {(out == ref_full).mean():.4f}, mean searched fraction {searched:.3f}",
          flush=True)


@pytest.mark.parametrize("algo,sub,scale,fmt,size", [(cs.ALGO_AW1, 5, 2, "i16", (400, 250, 128)),
                                                   (cs.ALGO_AW2, 5, 4, "ref", (400, 250, 128)),
                                                   (cs.ALGO_AW1, 3, 1, "ref", (400, 250, 128)),
                                                   (cs.ALGO_AW2, 3, 2, "i16", (400, 250, 128)),
                                                   (cs.ALGO_AW1, 5, 3, "i16", (333, 190, 64)),
                                                   (cs.ALGO_AW2, 3, 3, "ref", (130, 97, 8))])
def test_aw_band_search_matches_oracle(ctx, algo, sub, scale, fmt, size):
    """AW1 / AW2 around an initial map (MatcherUseDisparity of BlockMatcherAW1 / AW2): the
    oracle's adaptive-window matcher restricted to the same per-pixel bins, bit-exact;
    ragged sizes and the smallest range (D = 8) included."""
    W, H, D = size
    print(f"aw band search algo {algo} sub {sub} scale {scale} init {fmt} size {size}", flush=True)
    L, R = cs.synth_pair(97, W, H, 2, 6, D - 1, min(8, D - 1))
    full = cs.Matcher(ctx, W, H, algorithm=algo, window=(sub, sub), num_disp=D, tol=1)
    ref_full = full.run(L, R)
    rng = np.random.default_rng(20 + scale)
    small = ref_full[::scale, ::scale][: H // scale, : W // scale] // scale
    small = np.where(rng.random(small.shape) < 0.1, -1, small).astype(np.int16)
    small[: 60 // scale, : 130 // scale] = -1  # tiles with no valid value search the full range
    if fmt == "ref":
        init = np.zeros(small.shape + (4,), np.uint16)
        init[..., 0] = np.where(small >= 0, small + 1, 0xFFFD)
    else:
        init = small
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=(sub, sub), num_disp=D, tol=1, use_disp_input=True)
    got = m.run_init(L, R, init, maps=True)
    bins = _tile_bins(init, W, H, D, ref_format=(fmt == "ref"))
    mode = 1 if algo == cs.ALGO_AW1 else 2
    want = orc.aw_bins(mode, L, R, D, bins, sub=sub, tol=1, nthreads=8)[:3]
    for a, b in zip(got, want):
        assert np.array_equal(a, b), int((a != b).sum())
    searched = np.mean([bin(int(b)).count("1") for b in np.unique(bins)]) * 8 / D
    print(f"  agreement with the full range {(got[0] == ref_full).mean():.4f}, mean searched fraction {searched:.3f}",
          flush=True)


def test_band_search_without_valid_init_is_full_range(ctx):
    W, H, D = 320, 224, 64
    L, R = cs.synth_pair(92, W, H, 0, 6, D - 1, 8)
    init = np.full((H // 2, W // 2), -1, np.int16)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, num_disp=D, use_disp_input=True)
    got = m.run_init(L, R, init, maps=True)
    want = m.run(L, R, maps=True)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_band_search_configuration_errors(ctx):
    W, H, D = 256, 128, 64
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_SAD, window=(9, 9), num_disp=D)
    L, R = cs.synth_pair(93, W, H, 0, 4, D - 1, 8)
    with pytest.raises(cs.StereoError) as e:
        m.run_init(L, R, np.zeros((H // 2, W // 2), np.int16))
    assert e.value.status == cs.ERROR_INVALID_CONFIG


@pytest.mark.parametrize("fmt,size", [("i16", (331, 197)), ("ref", (256, 160)), ("i16", (1920, 1080))])
def test_disp_mipmaps_match_oracle(ctx, fmt, size):
    W, H = size
    print(f"mip maps {fmt} {W}x{H}", flush=True)
    rng = np.random.default_rng(W)
    d = rng.integers(0, 250, (H, W)).astype(np.int16)
    d[rng.random(d.shape) < 0.25] = -1
    if fmt == "ref":
        arr = np.zeros((H, W, 4), np.uint16)
        arr[..., 0] = np.where(d >= 0, d + 1, np.where(rng.random(d.shape) < 0.5, 0, 0xFFFD))
        img = cs.DeviceImage.from_array(ctx, arr, cs.PIXEL_RGBA_U16)
    else:
        arr = d
        img = cs.DeviceImage.from_array(ctx, arr, cs.PIXEL_LUM_I16)
    n = 5
    got = [lv.download() for lv in cs.disp_mipmaps(ctx, img, n)]
    want = orc.disp_mipmaps(arr, n)
    for k, (a, b) in enumerate(zip(got, want)):
        assert a.shape == b.shape, k
        assert np.array_equal(a, b), k


def test_disp_mipmaps_too_many_levels(ctx):
    img = cs.DeviceImage.from_array(ctx, np.zeros((10, 10), np.int16), cs.PIXEL_LUM_I16)
    with pytest.raises(cs.StereoError) as e:
        cs.disp_mipmaps(ctx, img, 4)  # 10 -> 4 -> 1 -> below one pixel
    assert e.value.status == cs.ERROR_INVALID_CONFIG
