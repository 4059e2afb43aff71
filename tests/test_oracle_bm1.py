"""The oracle's BlockMatcher1 emulation (BlockMatcher1.cu) pinned by known answers and by an
independent numpy restatement: per-shift similarity box sums from cumulative sums, the
reference's per-block evaluation lists, and a peak tracker vectorised over a block's
results."""
import numpy as np
import pytest

import orc

FIRST, LAST = 1, 0xFFEF
UNKNOWN, CANNOT, SATURATED, NOT_FOUND, NOT_SPECIFIC = 0, 0xFFFF, 0xFFFE, 0xFFFD, 0xFFFC
f32 = np.float32


def _sim(a, b, sim_type):
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    if sim_type == 0:
        return (255 - np.abs(a - b)) >> 1
    fa = f32(2.0) * (a.astype(f32) / f32(255.0) - f32(0.5))
    fb = f32(2.0) * (b.astype(f32) / f32(255.0) - f32(0.5))
    return ((fa * fb) * f32(127.0)).astype(np.int64)  # truncation toward zero


def _box(img, P):
    """P x P sums with the window's top-left at (y, x); zero beyond the image."""
    H, W = img.shape
    z = np.zeros((H + P + 1, W + P + 1), np.int64)
    z[1:H + 1, 1:W + 1] = img
    c = z.cumsum(0).cumsum(1)
    return c[P:P + H, P:P + W] - c[0:H, P:P + W] - c[P:P + H, 0:W] + c[0:H, 0:W]


def _f(v, P):
    return (v.astype(f32) / f32(P * P)) * (f32(100.0) / f32(127.0))


def _u16(f):
    t = np.clip((f / f32(100.0) + f32(1.0)) / f32(2.0), f32(0), f32(1))
    return (t * f32(65535.0)).astype(np.uint16)


def numpy_bm1(L, R, P, sim_type, dmin, dmax, off, ltor, min_delta, sim_thr, grad_thr, init=None, fill=0):
    H, W = L.shape
    CG = (32 - P + 1) // 8
    PCX = 8 * CG
    nbx = (W - (33 - PCX)) // PCX  # full blocks only
    nby = (H - (P - 1)) // 4
    out = np.full((H, W, 4), fill, np.uint16)
    boxes = {}

    def box_k(k):  # similarity box sums of L(x) against R(x + k)
        if k not in boxes:
            S = np.zeros((H, W), np.int64)
            xs = np.arange(W)
            ok = (xs + k >= 0) & (xs + k < W)
            S[:, ok] = _sim(L[:, ok], R[:, xs[ok] + k], sim_type)
            boxes[k] = _box(S, P)
        return boxes[k]

    Ls = _box(L.astype(np.int64), P)
    Lq = _box(L.astype(np.int64) ** 2, P)
    iy, ix = np.mgrid[0:4, 0:PCX]
    for by in range(nby):
        for bx in range(nbx):
            X0, Y0 = bx * PCX, by * 4
            if X0 + P > W or Y0 + P > H:
                continue
            wr = (X0 + ix + P < W) & (Y0 + iy + P < H)
            ysl, xsl = slice(Y0, Y0 + 4), slice(X0, X0 + PCX)
            disp = np.zeros((4, PCX), np.int64)
            if init is None:
                ssum = np.zeros((4, PCX), np.int64)
                ssq = np.zeros((4, PCX), np.int64)
                hh, ww = Ls[ysl, xsl].shape
                ssum[:hh, :ww] = Ls[ysl, xsl]
                ssq[:hh, :ww] = Lq[ysl, xsl]
                m = ssum.astype(f32) / f32(P * P)
                m2 = ssq.astype(f32) / f32(P * P)
                sd = np.sqrt(np.abs(m2 - m * m)) / f32(255.0) * f32(100.0)
                disp[sd < f32(grad_thr)] = CANNOT
                maxd = X0 - off + 1 if ltor else W - X0 - off
                starts, stop, cap = range(dmin, min(dmax, maxd), 16), None, 16
            else:
                iH, iW = init.shape[:2]
                s = W // iW
                lo_t, hi_t = [], []
                for t in range(32):
                    tl, th = LAST, FIRST
                    for r in range(CG):
                        cx, cy = (t // 4) * CG + r, t % 4
                        if not wr[cy, cx]:
                            continue
                        px, py = (X0 + cx + P // 2) // s, (Y0 + cy + P // 2) // s
                        idv = int(init[py, px, 0]) if (px < iW and py < iH) else 0
                        if idv <= LAST:
                            v = (s * idv) & 0xFFFF
                            tl, th = min(tl, v), max(th, v)
                            disp[cy, cx] = v
                        else:
                            disp[cy, cx] = idv
                    lo_t.append(tl)
                    hi_t.append(th)
                vl = [v for v in lo_t if FIRST <= v <= LAST]
                vh = [v for v in hi_t if FIRST <= v <= LAST]
                if not vl or not vh or min(vl) > max(vh):
                    # no valid range: every written result keeps its initial code
                    for cy, cx in zip(*np.nonzero(wr)):
                        out[Y0 + cy + P // 2, X0 + cx + P // 2] = (disp[cy, cx], 0, 0, 0)
                    continue
                lo, hi = min(vl), max(vh)
                d0 = 0 if lo < 4 else lo - 4
                starts, cap = range(d0, hi + 4 + 1, 16), min(16, hi + 4 - d0 + 1)
            ev = []
            for D0 in starts:
                tx = (X0 - D0 - off - 15) if ltor else (X0 + D0 + off)
                if tx >= 0 and tx + 48 <= W and Y0 + P + 3 <= H:
                    ev += [(D0 + r, (tx + 15 - r if ltor else tx + r) - X0) for r in range(cap)]
                elif Y0 + P + 3 <= H:
                    n, start = (W - tx - 32, tx) if tx >= 0 else (tx + 16, 0)
                    ev += [(D0 + r, (start + n - 1 - r if ltor else start + r) - X0) for r in range(max(n, 0))]
            mx = np.full((4, PCX), -32767, np.int64)
            prev2, prev, nxt, nxt2, last = (np.zeros((4, PCX), np.int64) for _ in range(5))
            for d, k in ev:
                v = box_k(k)[Y0:Y0 + 4, X0:X0 + PCX]
                if v.shape != (4, PCX):
                    v = np.pad(v, ((0, 4 - v.shape[0]), (0, PCX - v.shape[1])))
                act = disp <= LAST
                above = act & (v > mx)
                b1 = act & ~above & (disp + 1 == d + FIRST)
                b2 = act & ~above & (disp + 2 == d + FIRST)
                prev2 = np.where(above, prev, prev2)
                prev = np.where(above, last, prev)
                mx = np.where(above, v, mx)
                nxt = np.where(above | b1, v, nxt)
                nxt2 = np.where(above | b2, v, nxt2)
                disp = np.where(above, d + FIRST, disp)
                last = np.where(act, v, last)
            fm = _f(mx, P)
            dp = np.maximum(np.abs(_f(prev, P) - fm), np.abs(_f(prev2, P) - fm))
            dn = np.maximum(np.abs(_f(nxt, P) - fm), np.abs(_f(nxt2, P) - fm))
            spec = (dp >= f32(min_delta)) & (dn >= f32(min_delta))
            abv = fm >= f32(sim_thr)
            sat = (disp <= FIRST + dmin + 1) | (disp >= FIRST + dmax - 1)
            code = np.where(abv, np.where(sat, SATURATED, np.where(spec, disp, NOT_SPECIFIC)), NOT_FOUND)
            res = np.stack([code, _u16(fm), _u16(dp), _u16(dn)], -1).astype(np.uint16)
            special = (disp > LAST) | (disp == UNKNOWN)
            res[special] = 0
            res[..., 0][special] = disp[special]
            for cy in range(4):
                for cx in range(PCX):
                    if wr[cy, cx]:
                        out[Y0 + cy + P // 2, X0 + cx + P // 2] = res[cy, cx]
    return out


def _textured(seed, W, H):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (H, W)).astype(np.uint8)


def _scene(seed, W, H, shifts=(5, 13), ltor=True, noise=3):
    """Smooth texture (noise box-blurred 3x3) with flat patches; the left part of the
    image at the first disparity, the right part at the second, then noise."""
    rng = np.random.default_rng(seed)
    n = rng.integers(0, 256, (H + 2, W + 2)).astype(np.int64)
    L = sum(n[j:j + H, i:i + W] for j in range(3) for i in range(3)) // 9
    L = ((L - L.min()) * 255 // max(1, L.max() - L.min())).astype(np.int64)
    L[H // 4:H // 4 + 18, W // 4:W // 4 + 40] = 120  # flat: CannotEvaluate
    R = np.empty_like(L)
    split = W // 2
    for x in range(W):
        s = shifts[0] if x < split else shifts[1]
        src = x + s if ltor else x - s
        R[:, x] = L[:, min(max(src, 0), W - 1)]
    R = R + rng.integers(-noise, noise + 1, R.shape)
    return L.astype(np.uint8), np.clip(R, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("P,sim_type,ltor,dmin,dmax,off", [(5, 0, True, 2, 40, 1), (9, 0, True, 2, 40, 1),
                                                            (11, 1, True, 2, 40, 1), (15, 0, False, 2, 40, 1),
                                                            (9, 1, False, 2, 40, 1), (9, 0, True, 0, 64, 0),
                                                            (5, 1, True, 0, 64, 0), (11, 0, False, 0, 48, 0)])
def test_matches_numpy_restatement(P, sim_type, ltor, dmin, dmax, off):
    """With offset 0 and dmin 0, block origins 48, 96 (P = 9, 11) and 48, 96 (P = 5) meet a
    group start at the MaxDisparity bound."""
    W, H = 131, 47
    L, R = _scene(P, W, H, ltor=ltor)
    thr = (2.0, 70.0, 4.0) if sim_type == 0 else (0.5, 5.0, 4.0)
    args = dict(P=P, sim_type=sim_type, dmin=dmin, dmax=dmax, offset=off, ltor=ltor, min_delta=thr[0],
                sim_thresh=thr[1], grad_thresh=thr[2])
    got = orc.bm1(L, R, fill=0x1234, **args)
    want = numpy_bm1(L, R, P, sim_type, dmin, dmax, off, ltor, *thr, fill=0x1234)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())
    # the case exercises the codes: disparities of both regions and at least two of the
    # other outcomes
    c = got[..., 0]
    ds = set(np.unique(c[(c >= FIRST) & (c <= LAST)]) - FIRST)
    assert {5 - off, 13 - off} <= ds
    assert len({CANNOT, NOT_SPECIFIC, NOT_FOUND, SATURATED, UNKNOWN} & set(np.unique(c))) >= 2


@pytest.mark.parametrize("P", [5, 9])
def test_init_map_matches_numpy_restatement(P):
    W, H = 140, 52
    L = _textured(10 + P, W, H)
    R = np.roll(L, -9, axis=1)
    rng = np.random.default_rng(P)
    init = np.zeros((H // 2, W // 2, 4), np.uint16)
    init[..., 0] = FIRST + rng.integers(3, 6, init.shape[:2])
    init[rng.random(init.shape[:2]) < 0.1, 0] = NOT_FOUND
    init[:6, :20, 0] = 0  # a block whose threads see Unknown
    args = dict(P=P, sim_type=0, dmin=0, dmax=64, offset=0, ltor=True, min_delta=0.0, sim_thresh=0.0,
                grad_thresh=0.0)
    got = orc.bm1(L, R, init=init, fill=0x4321, **args)
    want = numpy_bm1(L, R, P, 0, 0, 64, 0, True, 0.0, 0.0, 0.0, init=init, fill=0x4321)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())


def test_known_shift_found_with_full_similarity():
    """R(x) = L(x + s): interior results of AbsoluteDifference find d = s, where every pixel
    pair is equal, so the similarity maximum is 127 per pixel (channel 1 = 65535)."""
    W, H, P, s = 160, 60, 9, 11
    L = _textured(3, W, H)
    R = np.zeros_like(L)
    R[:, :W - s] = L[:, s:]
    out = orc.bm1(L, R, P=P, dmin=0, dmax=32, min_delta=0.0, sim_thresh=0.0, grad_thresh=0.0)
    core = out[20:40, 60:120]
    assert (core[..., 0] == FIRST + s).all()
    assert (core[..., 1] == 65535).all()


def test_coverage_and_unwritten_pixels():
    """Written pixels are the centres of origins x <= min(W - P - 1, last block's end - 1)
    and y <= H - P - 1; everything else keeps its value."""
    W, H, P = 105, 40, 5  # 105 - 9 = 96 = 4 blocks of 24 origins: origins 96..99 are never covered
    L = _textured(4, W, H)
    out = orc.bm1(L, L, P=P, dmin=0, dmax=16, fill=0x7777)
    written = (out != 0x7777).any(-1)
    ys, xs = np.nonzero(written)
    assert xs.min() == P // 2 and ys.min() == P // 2
    assert xs.max() == 95 + P // 2
    assert ys.max() == H - P - 1 + P // 2


def test_only_full_blocks_are_launched():
    """EvalThreadConfig_ReadDepBlockSize keeps only full blocks (bOnlyFullBlocks defaults to
    true, KernelDriverBase.h:117-121): the remainder strips at the right and the bottom are
    never written."""
    W, H, P = 118, 47, 5  # 118 - 9 = 109: 4 full blocks of 24 origins; 47 - 4 = 43: 10 rows of 4
    L = _textured(9, W, H)
    out = orc.bm1(L, L, P=P, dmin=0, dmax=16, fill=0x7777)
    written = (out != 0x7777).any(-1)
    ys, xs = np.nonzero(written)
    assert xs.max() == 95 + P // 2  # origins 96..112 lie in the partial fifth block
    assert ys.max() == 39 + P // 2  # origin rows 40..41 lie in the partial eleventh block row
    assert np.array_equal(out, numpy_bm1(L, L, P, 0, 0, 16, 0, 1, 0.0, 0.0, 0.0, fill=0x7777))


def test_bottom_blocks_are_not_evaluated():
    """A block whose P + 3 read rows pass the bottom evaluates nothing: Unknown."""
    W, H, P = 96, 30, 9  # blocks start at rows 0, 4, .., 20; rows 20 + 12 > 30
    L = _textured(5, W, H)
    out = orc.bm1(L, L, P=P, dmin=0, dmax=16, grad_thresh=0.0)
    row = 20 + P // 2  # first result row of the block at Y0 = 20
    assert (out[row, P // 2:W - P - 1 + P // 2 - 8, 0] == UNKNOWN).all()
    assert (out[16 + P // 2, 40:60, 0] != UNKNOWN).all()


def test_flat_patches_cannot_be_evaluated():
    W, H = 105, 40  # 105 - 9 = 96: four full blocks, centres 2..97
    L = np.full((H, W), 90, np.uint8)
    out = orc.bm1(L, L, P=5, dmin=0, dmax=16, grad_thresh=0.5)
    written = out[..., 0][2:H - 4, 2:W - 7]
    assert (written == CANNOT).all()
    assert (out[2:H - 4, 2:W - 7, 1:] == 0).all()


def test_saturated_at_the_range_ends():
    W, H, P = 200, 40, 5
    L = _textured(6, W, H)
    out = orc.bm1(L, L, P=P, dmin=0, dmax=40)  # identical images: d = 0 = dmin
    assert (out[10:20, 60:150, 0] == SATURATED).all()


def test_signed_product_values():
    a = np.array([255, 0, 255, 128, 127, 200])
    b = np.array([255, 0, 0, 128, 127, 60])
    v = _sim(a, b, 1)
    assert v[0] == 127 and v[1] == 127 and v[2] == -127
    # the sign of the product of the centred values
    assert np.array_equal(np.sign(v[3:]), np.sign((a[3:] - 127.5) * (b[3:] - 127.5)).astype(int) * (v[3:] != 0))


def test_absolute_difference_identity_used_by_the_kernel():
    """(255 - |a - b|) >> 1 == 127 - (|a - b| >> 1) for all byte pairs: the GPU kernel sums
    |a - b| >> 1 (one v_sad_u8 and a shift) and subtracts from 127 P."""
    a, b = np.meshgrid(np.arange(256), np.arange(256))
    x = np.abs(a - b)
    assert np.array_equal((255 - x) >> 1, 127 - (x >> 1))
    assert not np.array_equal((255 - x) >> 1, 127 - ((x + 1) >> 1))
