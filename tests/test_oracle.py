"""Pins the CPU oracle (the checker of the HIP engine) without a GPU.

The reference holds no tests, fixtures or golden vectors for this path (its only
executable, CluTec.Cuda.UserTest1/UsertTest1.cpp, never runs a matcher), so the
oracle is pinned by closed forms, invariants, known answers and brute force on
small inputs, each chosen so that a dropped term, a wrong sign or index or a
transposed operand fails.
"""
import numpy as np
import pytest

import orc


def np_census(img, cw, ch):
    """Independent census: vectorised shifted comparisons."""
    img = img.astype(np.int32)
    H, W = img.shape
    rx, ry, n = cw // 2, ch // 2, cw * ch
    out = np.zeros((H, W), np.uint64)
    core = img[ry:H - ry, rx:W - rx]
    acc = np.zeros(core.shape, np.uint64)
    k = 0
    for dy in range(-ry, ry + 1):
        for dx in range(-rx, rx + 1):
            nb = img[ry + dy:H - ry + dy, rx + dx:W - rx + dx]
            acc |= (nb > core).astype(np.uint64) << np.uint64(n - 1 - k)
            k += 1
    out[ry:H - ry, rx:W - rx] = acc
    return out


# ---------------------------------------------------------------- census

def test_census_constant_image_is_zero():
    img = np.full((20, 24), 77, np.uint8)
    assert not orc.census(img, 9, 7).any()
    assert not orc.census3x3_ref(img).any()


def test_census_ramp_sets_right_columns():
    # I = x strictly increasing: neighbour > centre exactly when dx > 0
    img = np.tile(np.arange(40, dtype=np.uint8), (16, 1))
    c = orc.census(img, 9, 7)
    expect = 0
    k = 0
    for dy in range(-3, 4):
        for dx in range(-4, 5):
            if dx > 0:
                expect |= 1 << (62 - k)
            k += 1
    assert int(c[8, 20]) == expect
    assert int(c[0, 20]) == 0 and int(c[8, 2]) == 0  # border pixels are 0
    # reference 3x3 layout: right column bits 0x20, 0x08, 0x01
    assert int(orc.census3x3_ref(img)[5, 5]) == 0x20 | 0x08 | 0x01


def test_census_impulse_single_bit():
    img = np.zeros((21, 21), np.uint8)
    img[10, 10] = 200
    c = orc.census(img, 9, 7)
    # pixel (x=12, y=9) sees the impulse at dx=-2, dy=+1: raster index (1+3)*9 + (-2+4) = 38
    assert int(c[9, 12]) == 1 << (62 - 38)
    assert bin(int(c[9, 12])).count("1") == 1
    r = orc.census3x3_ref(img)
    assert int(r[9, 9]) == 0x01 and int(r[11, 11]) == 0x80 and int(r[10, 9]) == 0x08 and int(r[9, 10]) == 0x02


@pytest.mark.parametrize("cw,ch", [(9, 7), (7, 9), (5, 5), (3, 3)])
def test_census_matches_independent(cw, ch):
    rng = np.random.default_rng(cw * 10 + ch)
    img = rng.integers(0, 256, (23, 31), dtype=np.uint8)
    img[5:9, 5:9] = 100  # ties must not set bits (strict >)
    assert np.array_equal(orc.census(img, cw, ch), np_census(img, cw, ch))


def test_census3x3_matches_reference_bit_table():
    # CensusTransform.cu:106-117: (-1,-1)=0x80 (0,-1)=0x40 (+1,-1)=0x20 (-1,0)=0x10
    # (+1,0)=0x08 (-1,+1)=0x04 (0,+1)=0x02 (+1,+1)=0x01
    table = {(-1, -1): 0x80, (0, -1): 0x40, (1, -1): 0x20, (-1, 0): 0x10, (1, 0): 0x08,
             (-1, 1): 0x04, (0, 1): 0x02, (1, 1): 0x01}
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (12, 13), dtype=np.uint8)
    got = orc.census3x3_ref(img)
    for y in range(1, 11):
        for x in range(1, 12):
            v = 0
            for (dx, dy), b in table.items():
                if img[y + dy, x + dx] > img[y, x]:
                    v |= b
            assert got[y, x] == v


# ---------------------------------------------------------------- block matcher

def np_cost_volume(kind, L, R, D, ww, wh, cw, ch):
    """Independent O(k^2) cost volume with the plain validity rules."""
    H, W = L.shape
    ax, ay = ww // 2, wh // 2
    bx, by = ax + cw // 2, ay + ch // 2
    vol = np.full((H, W, D), -1, np.int64)
    for d in range(D):
        if kind == orc.COST_SAD:
            c = np.abs(L[:, d:].astype(np.int64) - R[:, :W - d].astype(np.int64))
        else:
            x = np.bitwise_xor(L[:, d:], R[:, :W - d])
            c = np.zeros(x.shape, np.int64)
            for b in range(64):
                c += ((x >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
        # c[y, x-d] is the cost of left pixel x
        cs = np.zeros((H, W), np.int64)
        cs[:, d:] = c
        for y in range(by, H - by):
            for x in range(max(bx, bx + d), W - bx):
                vol[y, x, d] = cs[y - ay:y + ay + 1, x - ax:x + ax + 1].sum()
    return vol


def wta_from_volume(vol, tol):
    H, W, D = vol.shape
    big = np.iinfo(np.int64).max
    v = np.where(vol < 0, big, vol)
    dL = np.where(vol.max(axis=2) < 0, -1, np.argmin(v, axis=2)).astype(np.int16)
    dR = np.full((H, W), -1, np.int16)
    best = np.full((H, W), big, np.int64)
    for d in range(D):
        for x in range(W):
            for y in range(H):
                a = vol[y, x, d]
                if a >= 0 and x - d >= 0 and a < best[y, x - d]:
                    best[y, x - d] = a
                    dR[y, x - d] = d
    out = np.full((H, W), -1, np.int16)
    for y in range(H):
        for x in range(W):
            d = dL[y, x]
            if d >= 0 and abs(int(dR[y, x - d]) - int(d)) <= tol:
                out[y, x] = d
    return out, dL, dR


@pytest.mark.parametrize("kind", [orc.COST_SAD, orc.COST_CENSUS])
@pytest.mark.parametrize("win", [(9, 9), (5, 5), (7, 3)])
def test_bm_matches_numpy_volume(kind, win):
    rng = np.random.default_rng(100 + kind * 10 + win[0])
    H, W, D = 26, 40, 12
    L8 = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R8 = np.roll(L8, -3, axis=1)
    R8 = np.clip(R8.astype(int) + rng.integers(-6, 7, (H, W)), 0, 255).astype(np.uint8)
    if kind == orc.COST_SAD:
        L, R, cwin = L8, R8, (1, 1)
    else:
        L, R, cwin = orc.census(L8, 5, 3), orc.census(R8, 5, 3), (5, 3)
    vol = np_cost_volume(kind, L, R, D, win[0], win[1], cwin[0], cwin[1])
    ref = wta_from_volume(vol, 1)
    got = orc.bm(kind, L, R, D, window=win, census_win=cwin)
    brute = orc.bm_brute(kind, L, R, D, window=win, census_win=cwin, with_volume=True)
    assert np.array_equal(brute[3].astype(np.int64), vol)
    for a, b, c in zip(got, ref, brute[:3]):
        assert np.array_equal(a, b)
        assert np.array_equal(c, b)


@pytest.mark.parametrize("kind", [orc.COST_SAD, orc.COST_CENSUS])
def test_bm_bins_against_the_cost_volume(kind):
    """The bins-restricted SAD / census matcher (the checker of the band searches): with
    every bin set it is orc_bm; with a random mask per left pixel, the left view is the
    lowest allowed cost of the independent numpy volume, and the right pixel xr takes the
    lowest cost over the left pixels xr + d whose own mask allows d (lowest d on ties)."""
    rng = np.random.default_rng(60 + kind)
    H, W, D = 24, 72, 32
    L8 = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R8 = np.roll(L8, -6, axis=1)
    R8 = np.clip(R8.astype(int) + rng.integers(-9, 10, (H, W)), 0, 255).astype(np.uint8)
    if kind == orc.COST_SAD:
        L, R, cwin, win = L8, R8, (1, 1), (5, 5)
    else:
        L, R, cwin, win = orc.census(L8, 5, 3), orc.census(R8, 5, 3), (5, 3), (5, 3)
    full = orc.bm(kind, L, R, D, window=win, census_win=cwin)
    every = orc.bm_bins(kind, L, R, D, np.full((H, W), 0xF, np.uint32), window=win, census_win=cwin)
    for a, b in zip(full, every):
        assert np.array_equal(a, b)
    vol = np_cost_volume(kind, L, R, D, win[0], win[1], cwin[0], cwin[1])
    bins = rng.integers(0, 16, (H, W)).astype(np.uint32)
    out, dl, dr = orc.bm_bins(kind, L, R, D, bins, window=win, census_win=cwin, nthreads=3)
    d = np.arange(D)
    allowed = (vol >= 0) & (((bins[:, :, None] >> (d[None, None, :] >> 3)) & 1) == 1)
    big = np.iinfo(np.int64).max
    q = np.where(allowed, vol, big)
    want_l = np.where(allowed.any(-1), q.argmin(-1), -1)
    assert np.array_equal(dl, want_l)
    want_r = np.full((H, W), -1)
    best = np.full((H, W), big)
    for dd in range(D):
        xs = np.arange(W - dd)
        cand = q[:, xs + dd, dd]  # the mask of left pixel xr + d decides
        better = cand < best[:, xs]
        best[:, xs] = np.where(better, cand, best[:, xs])
        want_r[:, xs] = np.where(better, dd, want_r[:, xs])
    assert np.array_equal(dr, want_r)
    assert np.array_equal(out, orc.crosscheck_plain(dl, dr, 1))
    # a mask that differs between the views changes the right map: the rule is not vacuous
    shifted = np.roll(bins, 5, axis=1)
    assert not np.array_equal(orc.bm_bins(kind, L, R, D, shifted, window=win, census_win=cwin)[2], dr)
    assert (dl >= 0).mean() > 0.5


def test_bm_noise_free_shift_is_exact():
    # R(x) = L(x + d*): the true disparity has cost 0, and textured windows are unique
    rng = np.random.default_rng(3)
    H, W, dstar = 40, 90, 11
    L = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R = np.zeros_like(L)
    R[:, :W - dstar] = L[:, dstar:]
    out, dL, dR, cost = orc.bm(orc.COST_SAD, L, R, 24, window=(5, 5), with_cost=True)
    inner = (slice(2, H - 2), slice(2 + 24, W - 2))
    assert (dL[inner] == dstar).all()
    assert (cost[inner] == 0).all()
    assert (out[inner] == dstar).all()
    # right map: right pixel xr matches left pixel xr + d*
    assert (dR[2:H - 2, 2:W - 2 - dstar - 5] == dstar).all()


def test_bm_ties_take_lowest_disparity():
    # a flat pair: every candidate costs 0 -> disparity 0 on both sides
    L = np.full((20, 40), 9, np.uint8)
    out, dL, dR = orc.bm(orc.COST_SAD, L, L, 16, window=(5, 5))
    v = dL >= 0
    assert v.any() and (dL[v] == 0).all() and (dR[dR >= 0] == 0).all()


def test_bm_threads_deterministic():
    import clutec_stereo as cs
    L, R = cs.synth_pair(5, 200, 120, 0, 6, 40, 8)
    a = orc.bm(orc.COST_SAD, L, R, 48, window=(9, 9), nthreads=1)
    b = orc.bm(orc.COST_SAD, L, R, 48, window=(9, 9), nthreads=8)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_crosscheck_plain_planted():
    dL = np.array([[-1, 0, 1, 2, 3, 3]], np.int16)
    dR = np.array([[0, 1, 3, 5, -1, 0]], np.int16)
    # x=1 d=0 -> dR[1]=1 ok(tol1); x=2 d=1 -> dR[1]=1 ok; x=3 d=2 -> dR[1]=1 ok;
    # x=4 d=3 -> dR[1]=1 |1-3|=2 fail; x=5 d=3 -> dR[2]=3 ok
    assert orc.crosscheck_plain(dL, dR, 1).tolist() == [[-1, 0, 1, 2, -1, 3]]
    assert orc.crosscheck_plain(dL, dR, 0).tolist() == [[-1, -1, 1, -1, -1, 3]]


def test_crosscheck_ref_codes_and_range():
    W = 8
    lr = np.zeros((1, W, 4), np.uint16)
    rl = np.zeros((1, W, 4), np.uint16)
    lr[0, :, 0] = [orc.DISP_UNKNOWN, 1, 2, 3, orc.DISP_NOT_FOUND, 10, 4, 3]
    lr[0, :, 1] = np.arange(W) + 100           # other channels travel with the pixel
    rl[0, :, 0] = [1, 1, 2, 9, 3, 3, 1, 2]
    out, mn, mx = orc.crosscheck_ref(lr, rl, thresh=1, ltor=1)
    # x=1 id1 d0 -> xr=1 id_R 1 keep; x=2 id2 d1 -> xr=1 id1 keep; x=3 id3 d2 -> xr=1 id1
    # |1-3| = 2 Inconsistent; x=4 NotFound passes; x=5 id10 d9 -> xr=-4 Inconsistent;
    # x=6 id4 d3 -> xr=3 id_R 9 Inconsistent; x=7 id3 d2 -> xr=5 id_R 3 keep
    assert out[0, :, 0].tolist() == [0, 1, 2, orc.DISP_INCONSISTENT, orc.DISP_NOT_FOUND, orc.DISP_INCONSISTENT,
                                     orc.DISP_INCONSISTENT, 3]
    assert out[0, :, 1].tolist() == lr[0, :, 1].tolist()
    assert (mn, mx) == (1, 3)
    # right-to-left direction maps x + d
    out2, _, _ = orc.crosscheck_ref(lr, rl, thresh=0, ltor=0)
    # x=1 id1 d0 -> xr=1 id1 keep; x=2 id2 d1 -> xr=3 id9 no; x=3 id3 d2 -> xr=5 id3 keep
    assert out2[0, 1, 0] == 1 and out2[0, 2, 0] == orc.DISP_INCONSISTENT and out2[0, 3, 0] == 3
    # nothing kept: the range stays at (Last, First)
    empty = np.zeros((1, 4, 4), np.uint16)
    _, mn, mx = orc.crosscheck_ref(empty, empty)
    assert (mn, mx) == (orc.DISP_LAST, orc.DISP_FIRST)


# ---------------------------------------------------------------- NCC

def np_ncc_volume(L, R, D, ww, wh):
    L = L.astype(np.float64)
    R = R.astype(np.float64)
    H, W = L.shape
    ax, ay = ww // 2, wh // 2
    N = ww * wh
    vol = np.full((H, W, D), np.nan)
    for y in range(ay, H - ay):
        for x in range(ax, W - ax):
            a = L[y - ay:y + ay + 1, x - ax:x + ax + 1].ravel()
            for d in range(D):
                if x - d < ax:
                    break
                b = R[y - ay:y + ay + 1, x - d - ax:x - d + ax + 1].ravel()
                va = N * (a * a).sum() - a.sum() ** 2
                vb = N * (b * b).sum() - b.sum() ** 2
                if va < N * N or vb < N * N:
                    vol[y, x, d] = 0.0
                else:
                    # textbook Pearson correlation of the two windows
                    vol[y, x, d] = np.corrcoef(a, b)[0, 1]
    return vol


def test_ncc_matches_pearson_and_gate():
    rng = np.random.default_rng(11)
    H, W, D = 16, 30, 9
    L = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R = np.roll(L, -2, axis=1)
    L[3:10, 3:12] = 50  # flat patch -> gated to 0
    out, dL, dR, sc, vol = orc.ncc(L, R, D, window=(5, 5), with_volume=True)
    ref = np_ncc_volume(L, R, D, 5, 5)
    m = ~np.isnan(ref)
    assert np.array_equal(m, ~np.isnan(vol))
    assert np.allclose(vol[m], ref[m], rtol=1e-12, atol=1e-12)
    # the left window at (x=7, y=5) lies inside the flat patch: every valid score is gated
    v57 = vol[5, 7]
    assert (~np.isnan(v57)).sum() == 6 and (v57[~np.isnan(v57)] == 0).all()


def test_ncc_affine_invariance_and_self():
    rng = np.random.default_rng(12)
    I = rng.integers(20, 100, (14, 24), dtype=np.uint8)
    _, _, _, _, v1 = orc.ncc(I, I, 6, window=(5, 5), with_volume=True)
    # NCC(I, I) at d=0 is 1
    good = ~np.isnan(v1[:, :, 0])
    assert np.allclose(v1[:, :, 0][good], 1.0)
    # NCC(2I+7, J) == NCC(I, J) (no clipping: 2*99+7 < 256)
    J = rng.integers(0, 256, (14, 24), dtype=np.uint8)
    _, _, _, _, va = orc.ncc(I, J, 6, window=(5, 5), with_volume=True)
    _, _, _, _, vb = orc.ncc((2 * I.astype(int) + 7).astype(np.uint8), J, 6, window=(5, 5), with_volume=True)
    m = ~np.isnan(va)
    assert np.allclose(va[m], vb[m], atol=1e-12)


# ---------------------------------------------------------------- adaptive windows

def np_aw_volume(mode, L, R, D, s):
    L = L.astype(np.int64)
    R = R.astype(np.int64)
    H, W = L.shape
    h, b = s // 2, s + s // 2
    vol = np.full((H, W, D), -1, np.int64)

    def S(y, x, d):
        a = L[y - h:y + h + 1, x - h:x + h + 1]
        c = R[y - h:y + h + 1, x - d - h:x - d + h + 1]
        ad = np.abs(a - c)
        return int((255 - ad).sum()) if mode == 1 else int(ad.sum())

    for y in range(b, H - b):
        for x in range(b, W - b):
            for d in range(D):
                if x - d < b:
                    break
                c = S(y, x, d)
                nb = sorted([S(y - s, x, d), S(y, x - s, d), S(y + s, x, d), S(y, x + s, d)])
                vol[y, x, d] = c + (nb[2] + nb[3] if mode == 1 else nb[0] + nb[1])
    return vol


@pytest.mark.parametrize("mode", [1, 2])
def test_aw_matches_numpy(mode):
    rng = np.random.default_rng(20 + mode)
    H, W, D = 24, 36, 8
    L = rng.integers(0, 256, (H, W), dtype=np.uint8)
    R = np.roll(L, -2, axis=1)
    out, dL, dR, val, vol = orc.aw(mode, L, R, D, sub=3, with_volume=True)
    ref = np_aw_volume(mode, L, R, D, 3)
    assert np.array_equal(vol.astype(np.int64), ref)


def test_aw1_aw2_equivalent_winners():
    import clutec_stereo as cs
    L, R = cs.synth_pair(2, 160, 96, 1, 4, 30, 8)
    a1 = orc.aw(1, L, R, 32, sub=5)
    a2 = orc.aw(2, L, R, 32, sub=5)
    for x, y in zip(a1[:3], a2[:3]):
        assert np.array_equal(x, y)
    v = a1[1] >= 0
    # similarity of the three chosen sub-patches = 255 * 3 * 25 - their SAD
    assert np.array_equal(a1[3][v], 255 * 75 - a2[3][v])


@pytest.mark.parametrize("mode", [1, 2])
def test_aw_bins_against_the_value_volume(mode):
    """The bins-restricted adaptive-window matcher: with every bin set it is orc_aw; with a
    random mask per left pixel, both views equal the best allowed value of the full volume
    (AW1 maximises, AW2 minimises; lowest d on ties)."""
    import numpy as np
    rng = np.random.default_rng(5 + mode)
    H, W, D = 34, 80, 24
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.roll(L, -4, axis=1)
    R = np.clip(R.astype(int) + rng.integers(-8, 9, R.shape), 0, 255).astype(np.uint8)
    full = orc.aw(mode, L, R, D, sub=3, tol=1)
    every = orc.aw_bins(mode, L, R, D, np.full((H, W), 0x7, np.uint32), sub=3, tol=1)
    for a, b in zip(full, every):
        assert np.array_equal(a, b)
    bins = rng.integers(0, 8, (H, W)).astype(np.uint32)
    out, dl, dr, val = orc.aw_bins(mode, L, R, D, bins, sub=3, tol=1)
    vol = orc.aw(mode, L, R, D, sub=3, tol=1, with_volume=True)[4].astype(np.int64)
    d = np.arange(D)
    allowed = (vol >= 0) & (((bins[:, :, None] >> (d[None, None, :] >> 3)) & 1) == 1)
    # a score to maximise, -1 where not allowed
    q = np.where(allowed, vol if mode == 1 else 100000 - vol, -1)
    want_l = np.where(allowed.any(-1), q.argmax(-1), -1)
    assert np.array_equal(dl, want_l)
    want_r = np.full((H, W), -1)
    best = np.full((H, W), -1)
    for dd in range(D):
        xs = np.arange(W - dd)
        cand = q[:, xs + dd, dd]
        better = cand > best[:, xs]
        best[:, xs] = np.where(better, cand, best[:, xs])
        want_r[:, xs] = np.where(better, dd, want_r[:, xs])
    assert np.array_equal(dr, want_r)
    assert np.array_equal(out, orc.crosscheck_plain(dl, dr, 1))
    yy, xx = np.nonzero(dl >= 0)
    assert np.array_equal(val[yy, xx], vol[yy, xx, dl[yy, xx]])
    assert (dl >= 0).mean() > 0.4


def test_ncc_bins_against_the_score_volume():
    """The bins-restricted NCC: with every bin set it is orc_ncc; with a random mask per left
    pixel, both views equal the argmax over the allowed candidates of the FP64 score volume
    (key floor(max(s, 0) * 2^23), lowest d on ties)."""
    import numpy as np
    rng = np.random.default_rng(4)
    H, W, D = 28, 90, 32
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.roll(L, -5, axis=1)
    R = np.clip(R.astype(int) + rng.integers(-6, 7, R.shape), 0, 255).astype(np.uint8)
    full = orc.ncc(L, R, D, window=(7, 7), tol=1)
    every = orc.ncc_bins(L, R, D, np.full((H, W), 0xF, np.uint32), window=(7, 7), tol=1)
    for a, b in zip(full, every):
        assert np.array_equal(np.nan_to_num(a, nan=-7), np.nan_to_num(b, nan=-7))
    bins = rng.integers(0, 16, (H, W)).astype(np.uint32)
    out, dl, dr, _ = orc.ncc_bins(L, R, D, bins, window=(7, 7), tol=1)
    vol = orc.ncc(L, R, D, window=(7, 7), tol=1, with_volume=True)[4]
    ok = ~np.isnan(vol)
    d = np.arange(D)
    allowed = ok & (((bins[:, :, None] >> (d[None, None, :] >> 3)) & 1) == 1)
    q = np.where(allowed, np.floor(np.maximum(np.nan_to_num(vol), 0.0) * 8388608.0), -1)
    want_l = np.where(allowed.any(-1), q.argmax(-1), -1)
    assert np.array_equal(dl, want_l)
    # right view: right pixel xr takes d over the left pixels xr + d that allow d
    want_r = np.full((H, W), -1)
    best = np.full((H, W), -1.0)
    for dd in range(D):
        xs = np.arange(W - dd)
        cand = q[:, xs + dd, dd]
        better = cand > best[:, xs]
        best[:, xs] = np.where(better, cand, best[:, xs])
        want_r[:, xs] = np.where(better, dd, want_r[:, xs])
    assert np.array_equal(dr, want_r)
    assert (dl >= 0).mean() > 0.5
