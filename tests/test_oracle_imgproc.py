"""The oracle's RectifyImage and ColorDisparityImage pinned by exact permutations, a float64
textbook evaluation of the OpenCV camera model, known palette / ramp values and numpy
restatements in float32."""
import numpy as np
import pytest

import orc

f32 = np.float32
FIRST, LAST = 1, 0xFFEF
CODES = [0, 0xFFFF, 0xFFFE, 0xFFFC, 0xFFFD, 0xFFFB, 0xFFFA]  # the palette's order


def _texture(seed, W, H):
    return np.random.default_rng(seed).integers(0, 256, (H, W)).astype(np.uint8)


# ---- rectification ------------------------------------------------------------------


def numpy_rectify(src, p, W, H):
    """float32, one operation at a time, in the order the camera model is written."""
    p = p.astype(f32)
    sH, sW = src.shape
    u, v = np.meshgrid(np.arange(W, dtype=f32), np.arange(H, dtype=f32))
    x = (u - p[11]) / p[9]
    y = (v - p[12]) / p[10]
    R = p[13:]
    X = R[0] * x + R[1] * y + R[2]
    Y = R[3] * x + R[4] * y + R[5]
    Z = R[6] * x + R[7] * y + R[8]
    xn, yn = X / Z, Y / Z
    r2 = xn * xn + yn * yn
    radial = f32(1) + r2 * (p[4] + r2 * (p[5] + r2 * p[8]))
    xy = xn * yn
    xd = xn * radial + f32(2) * p[6] * xy + p[7] * (r2 + f32(2) * xn * xn)
    yd = yn * radial + p[6] * (r2 + f32(2) * yn * yn) + f32(2) * p[7] * xy
    sx = p[0] * xd + p[2]
    sy = p[1] * yd + p[3]
    fx0, fy0 = np.floor(sx), np.floor(sy)
    nx, ny = np.floor(sx + f32(0.5)), np.floor(sy + f32(0.5))
    out = np.zeros((H, W), np.uint8)
    big = f32(2 ** 30)
    ok = (np.abs(fx0) < big) & (np.abs(fy0) < big)
    ix = np.where(ok, fx0, -2).astype(np.int64)
    iy = np.where(ok, fy0, -2).astype(np.int64)
    inside = (ix >= 0) & (iy >= 0) & (ix + 2 <= sW) & (iy + 2 <= sH)
    okn = (np.abs(nx) < big) & (np.abs(ny) < big)
    jx = np.where(okn, nx, -1).astype(np.int64)
    jy = np.where(okn, ny, -1).astype(np.int64)
    near = ~inside & (jx >= 0) & (jy >= 0) & (jx < sW) & (jy < sH)
    out[near] = src[jy[near], jx[near]]
    a, b = (sx - fx0)[inside], (sy - fy0)[inside]
    i0, j0 = ix[inside], iy[inside]
    n = [src[j0 + dy, i0 + dx].astype(f32) / f32(255) for dy, dx in ((0, 0), (0, 1), (1, 0), (1, 1))]
    f = (f32(1) - a) * (f32(1) - b) * n[0]
    f = f + a * (f32(1) - b) * n[1]
    f = f + (f32(1) - a) * b * n[2]
    f = f + a * b * n[3]
    out[inside] = (np.clip(f, f32(0), f32(1)) * f32(255) + f32(0.5)).astype(np.uint8)
    return out


def test_identity_is_exact():
    src = _texture(1, 97, 61)
    assert np.array_equal(orc.rectify(src, orc.rectify_params()), src)
    # realistic intrinsics: the round trip x = (u - c) / f, f x + c lands within an ulp of u
    p = orc.rectify_params(src=(712.5, 708.0, 48.3, 30.7))
    assert np.array_equal(orc.rectify(src, p), src)


def test_integer_shift_and_rotation_are_permutations():
    src = _texture(2, 80, 40)
    out = orc.rectify(src, orc.rectify_params(src=(1, 1, 0, 0), rect=(1, 1, 5, 0)))  # sx = u - 5
    assert np.array_equal(out[:, 5:], src[:, :-5]) and (out[:, :5] == 0).all()
    N = 41
    sq = _texture(3, N, N)
    c = (N - 1) / 2
    R = [[0, -1, 0], [1, 0, 0], [0, 0, 1]]  # (x, y) -> (-y, x)
    out = orc.rectify(sq, orc.rectify_params(src=(1, 1, c, c), R=R))
    v, u = np.mgrid[0:N, 0:N]
    assert np.array_equal(out, sq[u, 2 * int(c) - v])  # out(u, v) = src(2c - v, u)


def test_distortion_against_float64_camera_model():
    """A source holding its own x coordinate (and y in a second run) shows where each target
    pixel samples; compare with the OpenCV model evaluated in float64."""
    W, H = 256, 200
    dist = (-0.21, 0.07, 0.0012, -0.0008, -0.01)
    src_cam = (300.0, 290.0, 127.0, 99.0)
    rect_cam = (280.0, 280.0, 130.0, 100.0)
    ang = 0.03
    R = np.array([[np.cos(ang), -np.sin(ang), 0.01], [np.sin(ang), np.cos(ang), -0.02], [0.0, 0.015, 1.0]])
    p = orc.rectify_params(src_cam, dist, rect_cam, R)
    xs = np.tile(np.arange(W, dtype=np.uint8), (H, 1))
    ys = np.tile(np.minimum(np.arange(H), 255).astype(np.uint8)[:, None], (1, W))
    ox = orc.rectify(xs, p).astype(float)
    oy = orc.rectify(ys, p).astype(float)
    # float64 model
    u, v = np.meshgrid(np.arange(W, dtype=float), np.arange(H, dtype=float))
    Rf = p[13:].astype(float).reshape(3, 3)
    ray = np.stack([(u - rect_cam[2]) / rect_cam[0], (v - rect_cam[3]) / rect_cam[1], np.ones_like(u)])
    X, Y, Z = np.tensordot(Rf, ray, 1)
    xn, yn = X / Z, Y / Z
    k1, k2, p1, p2, k3 = dist
    r2 = xn ** 2 + yn ** 2
    rad = 1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3
    sx = src_cam[0] * (xn * rad + 2 * p1 * xn * yn + p2 * (r2 + 2 * xn ** 2)) + src_cam[2]
    sy = src_cam[1] * (yn * rad + p1 * (r2 + 2 * yn ** 2) + 2 * p2 * xn * yn) + src_cam[3]
    inner = (sx > 1) & (sx < W - 2) & (sy > 1) & (sy < H - 2)
    assert inner.mean() > 0.5
    assert np.abs(ox[inner] - sx[inner]).max() <= 0.51
    assert np.abs(oy[inner] - sy[inner]).max() <= 0.51
    # outside the source: 0
    far = (sx < -1) | (sx > W) | (sy < -1) | (sy > H)
    assert (ox[far] == 0).all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rectify_matches_numpy_restatement(seed):
    rng = np.random.default_rng(seed)
    sW, sH, W, H = 120, 90, 131, 77
    src = _texture(seed + 10, sW, sH)
    ang = rng.uniform(-0.05, 0.05, 3)
    cx, cy, cz = np.cos(ang)
    s_x, s_y, s_z = np.sin(ang)
    Rz = np.array([[cz, -s_z, 0], [s_z, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, s_y], [0, 1, 0], [-s_y, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -s_x], [0, s_x, cx]])
    p = orc.rectify_params((150 + rng.uniform(-5, 5), 148 + rng.uniform(-5, 5), 60 + rng.uniform(-3, 3), 45),
                           tuple(rng.uniform(-0.1, 0.1, 5) * [1, 0.3, 0.01, 0.01, 0.1]),
                           (140.0, 140.0, 65.0, 38.0), Rz @ Ry @ Rx)
    got = orc.rectify(src, p, W, H)
    want = numpy_rectify(src, p, W, H)
    assert np.array_equal(got, want), int((got != want).sum())
    assert 0.3 < (got > 0).mean()


# ---- disparity colouring --------------------------------------------------------------


def _norm(f, m):
    return (np.clip(np.nan_to_num(f, nan=0.0), f32(0), f32(1)) * f32(m) + f32(0.5)).astype(np.uint32)


def numpy_color(disp, style, dmin, dmax, off, colors):
    ids = disp[..., 0] if disp.ndim == 3 else disp
    v = ids.astype(f32)
    valid = (v >= f32(FIRST)) & (v <= f32(LAST))
    rng_ = f32(dmax - dmin)
    if ids.dtype == np.float32:
        fd = (v - f32(dmin)) / rng_
    else:
        fd = (ids.astype(np.int64) - dmin).astype(f32) / rng_
    H, W = ids.shape
    if style in (0, 2):
        out = np.zeros((H, W, 4), np.uint8)
        out[..., 3] = 255
        if style == 0:
            g = _norm(fd, 255).astype(np.uint8)
            out[..., 0] = out[..., 1] = out[..., 2] = np.where(valid, g, 0)
            other = ~valid
            out[other] = (255, 0, 0, 255)
            for k, c in enumerate(CODES):
                out[other & (v == f32(c))] = colors[k]
        else:
            with np.errstate(invalid="ignore", divide="ignore"):
                fB = f32(1) - np.minimum(fd / f32(0.7), f32(1))
                fG = f32(1) - np.minimum(np.abs(fd - f32(0.5)) / f32(0.59), f32(1))
                fR = f32(1) - np.minimum(np.abs(fd - f32(0.8)) / f32(0.7), f32(1))
                fM = np.fmax(np.fmax(fR, fG), fB)
                rgb = [_norm(c / fM, 255) for c in (fR, fG, fB)]
            for ch in range(3):
                out[..., ch] = np.where(valid, rgb[ch], 0)
        return out
    if style == 1:
        return np.where(valid, _norm(fd, 65535), 0).astype(np.uint16)
    if style == 3:
        return np.where(valid, v, 0).astype(np.uint16)
    if ids.dtype == np.float32:
        r = (f32(0x7FFF) + f32(off)) - (v - f32(FIRST))
        r = np.where(r > 0, np.minimum(r, f32(65535)), 0)
        return np.where(valid, np.trunc(r), 0).astype(np.uint16)
    return np.where(valid, (0x7FFF + off - (ids.astype(np.int64) - FIRST)) & 0xFFFF, 0).astype(np.uint16)


def _id_map(seed, W, H):
    rng = np.random.default_rng(seed)
    ids = rng.integers(FIRST, 300, (H, W)).astype(np.uint16)
    codes = np.array(CODES + [0xFFF0, 0xFFF5], np.uint16)
    pick = rng.random((H, W)) < 0.3
    ids[pick] = codes[rng.integers(0, len(codes), int(pick.sum()))]
    return ids


@pytest.mark.parametrize("style", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("fmt", ["rgba", "u16", "f32"])
def test_color_matches_numpy_restatement(style, fmt):
    ids = _id_map(style, 67, 31)
    if fmt == "rgba":
        disp = np.zeros(ids.shape + (4,), np.uint16)
        disp[..., 0] = ids
        disp[..., 1] = 999
    elif fmt == "u16":
        disp = ids
    else:
        disp = ids.astype(np.float32)
        disp[0, :5] = [1.5, 17.25, 0.5, 65519.5, 299.9]  # fractional and out-of-range floats
    colors = orc.DEFAULT_COLORS.copy()
    colors[6] = (1, 2, 3, 4)
    for dmin, dmax, off in [(1, 0xFFEF, 0), (20, 260, 7), (1, 300, 40000)]:
        got = orc.color_disparity(disp, style, dmin, dmax, off, colors)
        want = numpy_color(disp, style, dmin, dmax, off, colors)
        assert np.array_equal(got, want), (dmin, dmax, off)


def test_color_known_values():
    ids = np.array([[1, 150, 299, 0, 0xFFFF, 0xFFFA, 0xFFF0]], np.uint16)
    a = orc.color_disparity(ids, 0, 1, 299)
    assert list(a[0, 0]) == [0, 0, 0, 255] and list(a[0, 2]) == [255, 255, 255, 255]
    assert a[0, 1, 0] == 128  # 149 / 298 * 255 = 127.5, rounded half up
    assert list(a[0, 3]) == [10, 100, 180, 255] and list(a[0, 4]) == [0, 150, 50, 255]
    assert list(a[0, 5]) == [100, 10, 10, 255] and list(a[0, 6]) == [255, 0, 0, 255]
    g = orc.color_disparity(ids, 1, 1, 299)
    assert g[0, 0] == 0 and g[0, 2] == 65535 and g[0, 3] == 0
    assert list(orc.color_disparity(ids, 3)[0]) == [1, 150, 299, 0, 0, 0, 0]
    assert list(orc.color_disparity(ids, 4, inf_offset=5)[0, :3]) == [0x7FFF + 5, 0x7FFF + 5 - 149, 0x7FFF + 5 - 298]
    c = orc.color_disparity(ids, 2, 1, 299)
    # d = min: blue 1, green 1 - 0.5 / 0.59, red 0 -> (0, 39, 255)
    assert list(c[0, 0]) == [0, 39, 255, 255]
    assert list(c[0, 3]) == [0, 0, 0, 255]
