"""The oracle's disparity de-noise (Statistics.DispDenoise.cu) pinned by known answers and an
independent numpy restatement (float32 arithmetic, window moved by P at the borders)."""
import numpy as np
import pytest

import orc


def _numpy_denoise(d, P, ratio, cutoff, fill):
    H, W = d.shape
    out = np.empty_like(d)
    f = np.float32
    for y in range(H):
        ys = [yy + P if yy < 0 else (yy - P if yy >= H else yy) for yy in range(y - P // 2, y - P // 2 + P)]
        for x in range(W):
            xs = [xx + P if xx < 0 else (xx - P if xx >= W else xx) for xx in range(x - P // 2, x - P // 2 + P)]
            win = d[np.ix_(ys, xs)].astype(np.int64)
            ok = win >= 0
            cnt, s, s2 = int(ok.sum()), int(win[ok].sum()), int((win[ok] ** 2).sum())
            v = int(d[y, x])
            if not f(cnt) >= f(P * P) * f(ratio):
                r = -1
            elif v >= 0:
                mean = (f(s) - f(v)) / f(cnt - 1) if cnt > 1 else f(np.nan)
                r = int(np.floor(mean + f(0.5))) if abs(f(v) - mean) < f(cutoff) else -1
            elif fill:
                mean = f(s) / f(cnt)
                mean2 = f(s2) / f(cnt)
                sd = np.sqrt(np.abs(mean2 - mean * mean))
                r = int(np.floor(mean + f(0.5))) if sd < f(cutoff) else v
            else:
                r = v
            out[y, x] = r
    return out


def test_constant_map_is_unchanged():
    d = np.full((20, 24), 37, np.int16)
    assert np.array_equal(orc.disp_denoise(d, P=5), d)


def test_outlier_removed_neighbours_kept_hole_filled():
    d = np.full((15, 15), 50, np.int16)
    d[7, 7] = 60      # outlier: 10 away from the mean of its neighbours
    d[3, 3] = -1      # hole in a flat region
    out = orc.disp_denoise(d, P=5, mean_cutoff=1.0)
    assert out[7, 7] == -1
    assert out[7, 8] == 50 and out[6, 7] == 50
    assert out[3, 3] == 50
    # without fill the hole stays
    assert orc.disp_denoise(d, P=5, fill=False)[3, 3] == -1


def test_too_little_support_is_unknown():
    d = np.full((12, 12), -1, np.int16)
    d[::3, ::3] = 20  # 4 of 25 valid < 0.5 * 25
    out = orc.disp_denoise(d, P=5, support_ratio=0.5)
    assert (out == -1).all()


def test_reference_format_ids():
    ids = np.zeros((16, 16, 4), np.uint16)
    ids[..., 0] = 41
    ids[5, 5, 0] = 0xFFFA  # Inconsistent: filled from the flat window
    ids[9, 9, 0] = 49      # outlier
    out = orc.disp_denoise(ids, P=3)
    assert out.dtype == np.uint16
    assert out[5, 5] == 41 and out[9, 9] == 0 and out[0, 0] == 41


@pytest.mark.parametrize("P", [3, 4, 5, 16])
def test_matches_numpy_restatement(P):
    rng = np.random.default_rng(P)
    d = (30 + rng.integers(-2, 3, (19, 23))).astype(np.int16)
    d[rng.random(d.shape) < 0.2] = -1
    d[rng.random(d.shape) < 0.05] = 90
    for ratio, cutoff, fill in [(0.5, 1.0, True), (0.3, 2.0, False), (0.8, 0.5, True)]:
        got = orc.disp_denoise(d, P=P, support_ratio=ratio, mean_cutoff=cutoff, fill=fill)
        want = _numpy_denoise(d, P, ratio, cutoff, fill)
        assert np.array_equal(got, want), (ratio, cutoff, fill)


def _sd_limit2(c):
    """The kernel's fill-test limit: the square of the midpoint of c and the float below it."""
    c = np.float32(c)
    if not c > 0:
        return 0.0
    if np.isinf(c):
        return np.inf
    m = 0.5 * (float(np.nextafter(c, np.float32(0))) + float(c))
    return m * m


@pytest.mark.parametrize("c", [1.0, 0.75, 2.5, 1e-3, 3.0e4, np.float32(1.4e-45), 0.0, -1.0, np.inf, np.nan])
def test_sqrt_threshold_equivalence(c):
    """sqrt(x) < c, sqrt correctly rounded in float32, holds exactly when x < m*m in double:
    checked on floats around c*c (every float within 4096 ulps) and on random floats."""
    c32 = np.float32(c)
    lim = _sd_limit2(c32)
    with np.errstate(over="ignore", invalid="ignore"):
        centre = np.float32(c32 * c32) if np.isfinite(c32) and c32 > 0 else np.float32(1.0)
    bits = centre.view(np.int32) + np.arange(-4096, 4097, dtype=np.int32)
    bits = bits[(bits >= 0) & (bits < 0x7F800000)]
    rng = np.random.default_rng(0)
    xs = np.concatenate([bits.view(np.float32), rng.random(20000, dtype=np.float32) * np.float32(1e4),
                         np.array([0.0, np.nan], np.float32)])
    with np.errstate(invalid="ignore"):
        want = np.sqrt(xs) < c32
    got = xs.astype(np.float64) < lim
    assert np.array_equal(got, want)
