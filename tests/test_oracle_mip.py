"""The oracle's disparity mip maps (CreateDispMipMaps.cu) pinned by known answers and an
independent numpy restatement of the [1 2 1] passes."""
import numpy as np

import orc


def _numpy_next(src, unknown):
    """[1 2 1] along x, then along y with the factor 0.5; unknown values skipped."""
    def tap(a, b, c):
        va, vb, vc = a != unknown, b != unknown, c != unknown
        cnt = va * np.float32(1) + vb * np.float32(2) + vc * np.float32(1)
        s = np.where(va, a, np.float32(0))
        s = s + np.where(vb, b * np.float32(2), np.float32(0))
        s = s + np.where(vc, c, np.float32(0))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(cnt > 0, s / np.where(cnt > 0, cnt, 1).astype(np.float32), np.float32(unknown))

    H, W = src.shape
    Wn, Hn = (W - 1) // 2, (H - 1) // 2
    h = tap(src[:, 0:2 * Wn:2], src[:, 1:2 * Wn + 1:2], src[:, 2:2 * Wn + 2:2]).astype(np.float32)
    v = tap(h[0:2 * Hn:2], h[1:2 * Hn + 1:2], h[2:2 * Hn + 2:2]).astype(np.float32)
    return np.where(v != unknown, np.float32(0.5) * v, np.float32(unknown)).astype(np.float32)


def test_constant_map_halves_per_level():
    d = np.full((37, 50), 40, np.int16)
    lv = orc.disp_mipmaps(d, 4)
    assert [x.shape for x in lv] == [(37, 50), (18, 24), (8, 11), (3, 5)]
    for k, x in enumerate(lv):
        assert np.all(x == 40.0 / 2 ** k)


def test_unknown_is_skipped_and_propagates():
    d = np.full((9, 9), -1, np.int16)
    lv = orc.disp_mipmaps(d, 3)
    assert np.all(lv[1] == -1.0) and np.all(lv[2] == -1.0)
    # one row of taps: (x, 4, 8) with x unknown -> (2*4 + 8) / 3; then the vertical pass
    # over three equal rows gives the same value, halved
    d = np.full((3, 3), 0, np.int16)
    d[:, 0] = -1
    d[:, 1] = 4
    d[:, 2] = 8
    lv = orc.disp_mipmaps(d, 2)
    assert lv[1].shape == (1, 1)
    assert lv[1][0, 0] == np.float32(0.5) * (np.float32(16.0) / np.float32(3.0))


def test_reference_ids_outside_range_are_unknown():
    rgba = np.zeros((5, 6, 4), np.uint16)
    rgba[..., 0] = 7
    rgba[0, 0, 0] = 0xFFFA  # Inconsistent
    rgba[1, 1, 0] = 0xFFEF  # Last: valid
    lv = orc.disp_mipmaps(rgba, 1)[0]
    assert lv[0, 0] == 0.0 and lv[1, 1] == float(0xFFEF) and lv[2, 2] == 7.0


def test_matches_numpy_restatement():
    rng = np.random.default_rng(4)
    d = rng.integers(-1, 200, (83, 121)).astype(np.int16)
    d[rng.random(d.shape) < 0.3] = -1
    lv = orc.disp_mipmaps(d, 4)
    cur = np.where(d >= 0, d.astype(np.float32), np.float32(-1))
    assert np.array_equal(lv[0], cur)
    for k in range(1, 4):
        cur = _numpy_next(cur, -1.0)
        assert np.array_equal(lv[k], cur), k
