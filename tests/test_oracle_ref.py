"""The oracle's emulations of the reference NCC and AW1 operators (orc_ncc_ref,
orc_aw1_ref; BlockMatcherNCC.cu, BlockMatcherAW1.cu) pinned by known answers and by
independent numpy restatements built differently: window sums of every shift from
cumulative sums over the whole image, the reference's per-block evaluation lists, and the
peak tracker vectorised over a block's results.  Mistakes planted in the restatement (a
non-strict winner, a rounded byte score, the wrong neighbour rule, a partial remainder
block) change its output on the same scenes, so the comparisons can see them."""
import numpy as np
import pytest

import orc

FIRST, LAST = 1, 0xFFEF
UNKNOWN, CANNOT, SATURATED, NOT_FOUND, NOT_SPECIFIC = 0, 0xFFFF, 0xFFFE, 0xFFFD, 0xFFFC
f32 = np.float32


def _box(img, P):
    """P x P sums with the window's top-left at (y, x); zero beyond the image."""
    H, W = img.shape
    z = np.zeros((H + P + 1, W + P + 1), np.int64)
    z[1:H + 1, 1:W + 1] = img
    c = z.cumsum(0).cumsum(1)
    return c[P:P + H, P:P + W] - c[0:H, P:P + W] - c[P:P + H, 0:W] + c[0:H, 0:W]


def _shifted(R, k):
    """R(x + k) at x, zero where x + k is outside."""
    H, W = R.shape
    out = np.zeros((H, W), np.int64)
    xs = np.arange(W)
    ok = (xs + k >= 0) & (xs + k < W)
    out[:, ok] = R[:, xs[ok] + k]
    return out


def _eval_list(X0, Y0, W, H, tbx, tby, starts, cap, off, ltor):
    """(d, k): disparity d tests the right window starting k columns right of the left one."""
    ev = []
    for D0 in starts:
        tx = (X0 - D0 - off - 15) if ltor else (X0 + D0 + off)
        if tx >= 0 and tx + tbx + 15 <= W and Y0 + tby <= H:
            ev += [(D0 + r, (tx + 15 - r if ltor else tx + r) - X0) for r in range(cap)]
        elif Y0 + tby <= H:
            n, start = (W - tx - tbx + 1, tx) if tx >= 0 else (tx + 16, 0)
            ev += [(D0 + r, (start + n - 1 - r if ltor else start + r) - X0) for r in range(max(n, 0))]
    return ev


def _band(init, W, X0, Y0, half, wr, threads):
    """MatcherUseDisparity: initial values per result and the warp's valid min / max.
    threads: for each reference thread, its results as (row, col) in the block."""
    iH, iW = init.shape[:2]
    s = W // iW
    disp = {}
    lo_t, hi_t = [], []
    for res in threads:
        tl, th = LAST, FIRST
        for cy, cx in res:
            if not wr[cy, cx]:
                continue
            px, py = (X0 + cx + half) // s, (Y0 + cy + half) // s
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
        return disp, None
    return disp, (min(vl), max(vh))


def _track(ev, values, disp, mx0, strict=True):
    """The peak tracker over the evaluation list, vectorised over the block's results."""
    mx = np.full(disp.shape, mx0, np.int64)
    prev2, prev, nxt, nxt2, last = (np.zeros(disp.shape, np.int64) for _ in range(5))
    for d, k in ev:
        v = values(k)
        act = disp <= LAST
        above = act & ((v > mx) if strict else (v >= mx))
        b1 = act & ~above & (disp + 1 == d + FIRST)
        b2 = act & ~above & (disp + 2 == d + FIRST)
        prev2 = np.where(above, prev, prev2)
        prev = np.where(above, last, prev)
        mx = np.where(above, v, mx)
        nxt = np.where(above | b1, v, nxt)
        nxt2 = np.where(above | b2, v, nxt2)
        disp = np.where(above, d + FIRST, disp)
        last = np.where(act, v, last)
    return disp, mx, prev, prev2, nxt, nxt2


# ------------------------------------------------------------------------------ NCC

def numpy_ncc_ref(L, R, P, dmin, dmax, off, ltor, min_delta, ncc_thr, grad_thr, init=None, fill=0,
                  mistake=None):
    H, W = L.shape
    CG = (33 - P) // 2
    PCX, PCY, TBX, TBY = 2 * CG, 16, 33, 15 + P
    full_only = mistake != "partial_block"
    nbx = (W - (TBX - PCX)) // PCX + (0 if full_only else int((W - (TBX - PCX)) % PCX > 0))
    nby = (H - (TBY - PCY)) // PCY
    fN = f32(P * P)
    Li = L.astype(np.int64)
    Ri = R.astype(np.int64)
    gx = np.abs(Li - np.concatenate([Li[:, 1:], np.zeros((H, 1), np.int64)], 1))
    G = _box(gx, P)
    Ls, Lq = _box(Li, P), _box(Li * Li, P)
    Rs, Rq = _box(Ri, P), _box(Ri * Ri, P)
    cross = {}

    def cross_k(k):
        if k not in cross:
            cross[k] = _box(Li * _shifted(Ri, k), P)
        return cross[k]

    out = np.full((H, W, 4), fill, np.uint16)
    iy, ix = np.mgrid[0:PCY, 0:PCX]
    threads = [[(t % 16, (t // 16) * CG + r) for r in range(CG)] for t in range(32)]
    for by in range(nby):
        for bx in range(nbx):
            X0, Y0 = bx * PCX, by * PCY
            if X0 + P > W or Y0 + P > H:
                continue
            wr = (X0 + ix + P < W) & (Y0 + iy + P < H)

            def blk(a):
                v = np.zeros((PCY, PCX), a.dtype)
                t = a[Y0:Y0 + PCY, X0:X0 + PCX]
                v[:t.shape[0], :t.shape[1]] = t
                return v

            disp = np.zeros((PCY, PCX), np.int64)
            if init is None:
                disp[blk(G).astype(f32) / fN < f32(grad_thr)] = CANNOT
                if (disp == CANNOT).all():
                    for cy, cx in zip(*np.nonzero(wr)):
                        out[Y0 + cy + P // 2, X0 + cx + P // 2] = (CANNOT, 0, 0, 0)
                    continue
                maxd = X0 - off + 1 if ltor else W - X0 - off
                ev = _eval_list(X0, Y0, W, H, TBX, TBY, range(dmin, min(dmax, maxd), 16), 16, off, ltor)
            else:
                vals, rng_ = _band(init, W, X0, Y0, P // 2, wr, threads)
                for (cy, cx), v in vals.items():
                    disp[cy, cx] = v
                if rng_ is None:
                    for cy, cx in zip(*np.nonzero(wr)):
                        out[Y0 + cy + P // 2, X0 + cx + P // 2] = (disp[cy, cx], 0, 0, 0)
                    continue
                lo, hi = rng_
                d0 = 0 if lo < 4 else lo - 4
                ev = _eval_list(X0, Y0, W, H, TBX, TBY, range(d0, hi + 4 + 1, 16), min(16, hi + 4 - d0 + 1), off,
                                ltor)
            with np.errstate(invalid="ignore", divide="ignore"):
                bm = blk(Ls).astype(f32) / fN
                bsd = np.sqrt(blk(Lq).astype(f32) / fN - bm * bm)

                def score(k):
                    xs = np.clip(X0 + ix + k, 0, W - 1)
                    ys = np.clip(Y0 + iy, 0, H - 1)
                    tm = Rs[ys, xs].astype(f32) / fN
                    tsd = np.sqrt(Rq[ys, xs].astype(f32) / fN - tm * tm)
                    norm = bsd * tsd
                    c = blk(cross_k(k)).astype(f32) / fN
                    ncc = np.where(norm < f32(1.0), f32(0.0), (c - bm * tm) / norm)
                    ncc = np.fmax(f32(0.0), ncc)
                    q = np.fmin(f32(255.0), ncc * f32(255.0))
                    return np.rint(q).astype(np.int64) if mistake == "rounded_byte" else q.astype(np.int64)

                disp, mx, prev, prev2, nxt, nxt2 = _track(ev, score, disp, 0, strict=mistake != "non_strict")
            d1 = np.maximum(np.abs(prev - mx), np.abs(prev2 - mx))
            d2 = np.maximum(np.abs(nxt - mx), np.abs(nxt2 - mx))
            abv = (mx.astype(f32) / f32(255.0)) >= f32(ncc_thr)
            sat = (disp <= FIRST + dmin + 1) | (disp >= FIRST + dmax - 1)
            spec = (d1 >= min_delta) & (d2 >= min_delta)
            code = np.where(abv, np.where(sat, SATURATED, np.where(spec, disp, NOT_SPECIFIC)), NOT_FOUND)
            res = np.stack([code, mx, d1, d2], -1)
            special = (disp > LAST) | (disp == UNKNOWN)
            res[special] = 0
            res[..., 0][special] = disp[special]
            for cy, cx in zip(*np.nonzero(wr)):
                out[Y0 + cy + P // 2, X0 + cx + P // 2] = res[cy, cx]
    return out


def _scene(seed, W, H, shifts=(5, 13), ltor=True, noise=3):
    """Smooth texture with a flat patch; the left part of the image at the first
    disparity, the right part at the second, then noise."""
    rng = np.random.default_rng(seed)
    n = rng.integers(0, 256, (H + 2, W + 2)).astype(np.int64)
    L = sum(n[j:j + H, i:i + W] for j in range(3) for i in range(3)) // 9
    L = ((L - L.min()) * 255 // max(1, L.max() - L.min())).astype(np.int64)
    L[H // 4:H // 4 + 18, W // 4:W // 4 + 40] = 120
    R = np.empty_like(L)
    split = W // 2
    for x in range(W):
        s = shifts[0] if x < split else shifts[1]
        src = x + s if ltor else x - s
        R[:, x] = L[:, min(max(src, 0), W - 1)]
    R = R + rng.integers(-noise, noise + 1, R.shape)
    return L.astype(np.uint8), np.clip(R, 0, 255).astype(np.uint8)


NCC_CASES = [(11, True, 0, 48, 0, 2, 0.6, 3.0), (9, True, 2, 40, 1, 3, 0.5, 4.0), (5, True, 0, 64, 0, 1, 0.7, 2.0),
             (11, False, 2, 40, 1, 2, 0.6, 3.0), (9, False, 0, 48, 0, 4, 0.8, 5.0)]


@pytest.mark.parametrize("P,ltor,dmin,dmax,off,md,nt,gt", NCC_CASES)
def test_ncc_ref_matches_numpy_restatement(P, ltor, dmin, dmax, off, md, nt, gt):
    W, H = 141, 71
    L, R = _scene(P, W, H, ltor=ltor)
    got = orc.ncc_ref(L, R, P=P, dmin=dmin, dmax=dmax, offset=off, ltor=ltor, min_delta=md, ncc_thresh=nt,
                      grad_thresh=gt, fill=0x1234)
    want = numpy_ncc_ref(L, R, P, dmin, dmax, off, ltor, md, nt, gt, fill=0x1234)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())
    c = got[..., 0]
    ds = set(np.unique(c[(c >= FIRST) & (c <= LAST)]) - FIRST)
    assert {5 - off, 13 - off} <= ds
    assert len({CANNOT, NOT_SPECIFIC, NOT_FOUND, SATURATED, UNKNOWN} & set(np.unique(c))) >= 1


@pytest.mark.parametrize("mistake", ["non_strict", "rounded_byte", "partial_block"])
def test_ncc_ref_restatement_sees_planted_mistakes(mistake):
    W, H = 141, 71
    L, R = _scene(11, W, H)
    got = orc.ncc_ref(L, R, P=11, dmin=0, dmax=48, min_delta=2, ncc_thresh=0.6, grad_thresh=3.0, fill=0x1234)
    bad = numpy_ncc_ref(L, R, 11, 0, 48, 0, True, 2, 0.6, 3.0, fill=0x1234, mistake=mistake)
    assert not np.array_equal(got, bad)


@pytest.mark.parametrize("P,scale", [(11, 2), (5, 4), (9, 1)])
def test_ncc_ref_init_map_matches_numpy_restatement(P, scale):
    W, H = 150, 64
    L = _scene(20 + P, W, H)[0]
    R = np.roll(L, -9, axis=1)
    rng = np.random.default_rng(P)
    init = np.zeros((H // scale, W // scale, 4), np.uint16)
    init[..., 0] = FIRST + rng.integers(8, 11, init.shape[:2]) // scale
    init[rng.random(init.shape[:2]) < 0.1, 0] = NOT_FOUND
    init[:8 // scale + 1, :40 // scale, 0] = 0  # threads that see Unknown
    got = orc.ncc_ref(L, R, P=P, dmin=0, dmax=64, min_delta=1, ncc_thresh=0.3, init=init, fill=0x4321)
    want = numpy_ncc_ref(L, R, P, 0, 64, 0, True, 1, 0.3, 0.0, init=init, fill=0x4321)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())
    assert (got[..., 0] == FIRST + 9).mean() > 0.3


def test_ncc_ref_known_shift_scores_255():
    """R(x) = L(x + s) exactly: interior results find d = s with NCC 1, whose byte is 255 or,
    where the float formula rounds the ratio just below 1, 254; the diffs are the drops to
    the neighbouring disparities."""
    W, H, s = 160, 64, 7
    rng = np.random.default_rng(2)
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.zeros_like(L)
    R[:, :W - s] = L[:, s:]
    out = orc.ncc_ref(L, R, P=9, dmin=0, dmax=32)
    core = out[20:40, 60:120]
    assert (core[..., 0] == FIRST + s).all()
    assert (core[..., 1] >= 254).all() and (core[..., 1] == 255).any()
    assert (core[..., 2] > 100).all() and (core[..., 3] > 100).all()


def test_ncc_ref_byte_is_truncated_float_formula():
    """One window by hand: the byte is floor(255 max(0, ncc)) of the reference's float
    formula (cross mean minus product of means over the product of standard deviations)."""
    rng = np.random.default_rng(8)
    W, H, P = 64, 40, 5
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.clip(L.astype(int) + rng.integers(-40, 41, L.shape), 0, 255).astype(np.uint8)
    out = orc.ncc_ref(L, R, P=P, dmin=0, dmax=1, ltor=True)  # d = 0 only
    y, x = 10, 20
    a = L[y - 2:y + 3, x - 2:x + 3].astype(np.int64)
    b = R[y - 2:y + 3, x - 2:x + 3].astype(np.int64)
    n = f32(25)
    am, bm = f32(a.sum()) / n, f32(b.sum()) / n
    asd = np.sqrt(f32((a * a).sum()) / n - am * am)
    bsd = np.sqrt(f32((b * b).sum()) / n - bm * bm)
    ncc = (f32((a * b).sum()) / n - am * bm) / (asd * bsd)
    assert out[y, x, 1] == int(min(f32(255), max(f32(0), ncc) * f32(255)))
    exact = np.corrcoef(a.ravel(), b.ravel())[0, 1]
    assert abs(out[y, x, 1] - 255 * exact) < 1.01


def test_ncc_ref_flat_image_cannot_be_evaluated_and_written_area():
    """A flat image: every block stops at the gradient gate and writes CannotEvaluate over
    the centres of its written results; remainder strips stay untouched."""
    W, H, P = 120, 60, 11  # nbx = (120 - 11) // 22 = 4 (origins 0..87), nby = (60 - 10) // 16 = 3
    L = np.full((H, W), 77, np.uint8)
    out = orc.ncc_ref(L, L, P=P, dmin=0, dmax=16, grad_thresh=0.5, fill=0x7777)
    written = (out != 0x7777).any(-1)
    ys, xs = np.nonzero(written)
    assert xs.min() == P // 2 and ys.min() == P // 2
    assert xs.max() == 87 + P // 2 and ys.max() == 47 + P // 2
    assert (out[written][:, 0] == CANNOT).all() and (out[written][:, 1:] == 0).all()


def test_ncc_ref_saturated_at_range_ends():
    W, H = 160, 60
    L = _scene(3, W, H, shifts=(0, 0))[0]
    out = orc.ncc_ref(L, L, P=9, dmin=0, dmax=32)  # identical images: d = 0 = dmin
    assert (out[38:52, 90:140, 0] == SATURATED).all()
    # the flat patch scores 0 at every disparity: no byte exceeds the initial 0, so its
    # results stay Unknown
    assert (out[20:28, 50:70, 0] == UNKNOWN).all()


# ------------------------------------------------------------------------------ AW1

def _aw_sim(a, b, sim_type):
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    if sim_type == 0:
        return 255 - np.abs(a - b)
    fa = f32(2.0) * (a.astype(f32) / f32(255.0) - f32(0.5))
    fb = f32(2.0) * (b.astype(f32) / f32(255.0) - f32(0.5))
    return ((fa * fb) * f32(255.0)).astype(np.int64)


def _aw_f(v, ne):
    return (v.astype(f32) / f32(ne)) * (f32(100.0) / f32(255.0))


def _u16(f):
    t = np.clip((f / f32(100.0) + f32(1.0)) / f32(2.0), f32(0), f32(1))
    return (t * f32(65535.0)).astype(np.uint16)


def numpy_aw1_ref(L, R, s, sim_type, dmin, dmax, off, ltor, min_delta, sim_thr, grad_thr, init=None, fill=0,
                  mistake=None):
    H, W = L.shape
    CG = (33 - s) // 2
    P3 = 3 * s
    PCX, PCY, TBX = 2 * CG - 2 * s, 16 - 2 * s, 32
    TBY = PCY + P3 - 1
    NE = 3 * s * s
    nbx = (W - (TBX - PCX)) // PCX
    nby = (H - (TBY - PCY)) // PCY
    Li = L.astype(np.int64)
    Ls, Lq = _box(Li, P3), _box(Li * Li, P3)
    sub = {}

    def sub_k(k):  # s x s similarity sums of L(x) against R(x + k), top-left at (y, x)
        if k not in sub:
            xs = np.arange(W)
            ok = (xs + k >= 0) & (xs + k < W)
            S = np.zeros((H, W), np.int64)
            S[:, ok] = _aw_sim(L[:, ok], R[:, xs[ok] + k], sim_type)
            sub[k] = _box(S, s)
        return sub[k]

    out = np.full((H, W, 4), fill, np.uint16)
    iy, ix = np.mgrid[0:PCY, 0:PCX]
    NR = PCX * PCY
    threads = [[divmod(r * 32 + t, PCX) for r in range(-(-NR // 32)) if r * 32 + t < NR] for t in range(32)]
    for by in range(nby):
        for bx in range(nbx):
            X0, Y0 = bx * PCX, by * PCY
            if X0 + P3 > W or Y0 + P3 > H:
                continue
            wr = (X0 + ix + P3 < W) & (Y0 + iy + P3 < H)
            disp = np.zeros((PCY, PCX), np.int64)
            if init is None:
                m = Ls[Y0:Y0 + PCY, X0:X0 + PCX].astype(f32) / f32(P3 * P3)
                m2 = Lq[Y0:Y0 + PCY, X0:X0 + PCX].astype(f32) / f32(P3 * P3)
                sd = np.sqrt(np.abs(m2 - m * m)).astype(np.int64).astype(f32) / f32(255.0) * f32(100.0)
                disp[sd < f32(grad_thr)] = CANNOT
                if (disp == CANNOT).all():
                    for cy, cx in zip(*np.nonzero(wr)):
                        out[Y0 + cy + P3 // 2, X0 + cx + P3 // 2] = (CANNOT, 0, 0, 0)
                    continue
                maxd = X0 - off + 1 if ltor else W - X0 - off
                ev = _eval_list(X0, Y0, W, H, TBX, TBY, range(dmin, min(dmax, maxd), 16), 16, off, ltor)
            else:
                vals, rng_ = _band(init, W, X0, Y0, P3 // 2, wr, threads)
                for (cy, cx), v in vals.items():
                    disp[cy, cx] = v
                if rng_ is None:
                    for cy, cx in zip(*np.nonzero(wr)):
                        out[Y0 + cy + P3 // 2, X0 + cx + P3 // 2] = (disp[cy, cx], 0, 0, 0)
                    continue
                lo, hi = rng_
                d0 = 0 if lo < 4 else lo - 4
                ev = _eval_list(X0, Y0, W, H, TBX, TBY, range(d0, hi + 4 + 1, 16), min(16, hi + 4 - d0 + 1), off,
                                ltor)

            def value(k):
                S = sub_k(k)

                def at(dy, dx):
                    v = np.zeros((PCY, PCX), np.int64)
                    t = S[Y0 + dy:Y0 + dy + PCY, X0 + dx:X0 + dx + PCX]
                    v[:t.shape[0], :t.shape[1]] = t
                    return v

                c = at(s, s)
                nb = np.stack([at(0, s), at(s, 0), at(2 * s, s), at(s, 2 * s)])
                if mistake == "corners":  # the diagonal neighbours instead of the edge ones
                    nb = np.stack([at(0, 0), at(0, 2 * s), at(2 * s, 0), at(2 * s, 2 * s)])
                nb.sort(axis=0)
                return c + nb[3] + nb[2]

            disp, mx, prev, prev2, nxt, nxt2 = _track(ev, value, disp, -32767)
            fm = _aw_f(mx, NE)
            dp = np.maximum(np.abs(_aw_f(prev, NE) - fm), np.abs(_aw_f(prev2, NE) - fm))
            dn = np.maximum(np.abs(_aw_f(nxt, NE) - fm), np.abs(_aw_f(nxt2, NE) - fm))
            spec = (dp >= f32(min_delta)) & (dn >= f32(min_delta))
            abv = fm >= f32(sim_thr)
            sat = (disp <= FIRST + dmin + 1) | (disp >= FIRST + dmax - 1)
            code = np.where(abv, np.where(sat, SATURATED, np.where(spec, disp, NOT_SPECIFIC)), NOT_FOUND)
            res = np.stack([code, _u16(fm), _u16(dp), _u16(dn)], -1).astype(np.uint16)
            special = (disp > LAST) | (disp == UNKNOWN)
            res[special] = 0
            res[..., 0][special] = disp[special]
            for cy, cx in zip(*np.nonzero(wr)):
                out[Y0 + cy + P3 // 2, X0 + cx + P3 // 2] = res[cy, cx]
    return out


AW_CASES = [(5, 0, True, 0, 48, 0, 1.0, 80.0, 3.0), (3, 0, True, 2, 40, 1, 0.5, 75.0, 2.0),
            (5, 1, True, 0, 48, 0, 0.3, 5.0, 3.0), (3, 1, False, 2, 40, 1, 0.5, 5.0, 2.0),
            (5, 0, False, 0, 64, 0, 1.0, 80.0, 4.0)]


@pytest.mark.parametrize("s,sim,ltor,dmin,dmax,off,md,st,gt", AW_CASES)
def test_aw1_ref_matches_numpy_restatement(s, sim, ltor, dmin, dmax, off, md, st, gt):
    W, H = 143, 69
    L, R = _scene(30 + s, W, H, ltor=ltor)
    got = orc.aw1_ref(L, R, sub=s, sim_type=sim, dmin=dmin, dmax=dmax, offset=off, ltor=ltor, min_delta=md,
                      sim_thresh=st, grad_thresh=gt, fill=0x1234)
    want = numpy_aw1_ref(L, R, s, sim, dmin, dmax, off, ltor, md, st, gt, fill=0x1234)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())
    c = got[..., 0]
    ds = set(np.unique(c[(c >= FIRST) & (c <= LAST)]) - FIRST)
    assert {5 - off, 13 - off} <= ds
    assert len({CANNOT, NOT_SPECIFIC, NOT_FOUND, SATURATED, UNKNOWN} & set(np.unique(c))) >= 1


def test_aw1_ref_restatement_sees_planted_mistake():
    W, H = 143, 69
    L, R = _scene(35, W, H)
    got = orc.aw1_ref(L, R, sub=5, dmin=0, dmax=48, min_delta=1.0, sim_thresh=80.0, grad_thresh=3.0, fill=0x1234)
    bad = numpy_aw1_ref(L, R, 5, 0, 0, 48, 0, True, 1.0, 80.0, 3.0, fill=0x1234, mistake="corners")
    assert not np.array_equal(got, bad)


@pytest.mark.parametrize("s,scale", [(5, 2), (3, 4)])
def test_aw1_ref_init_map_matches_numpy_restatement(s, scale):
    W, H = 150, 64
    L = _scene(40 + s, W, H)[0]
    R = np.roll(L, -9, axis=1)
    rng = np.random.default_rng(s)
    init = np.zeros((H // scale, W // scale, 4), np.uint16)
    init[..., 0] = FIRST + rng.integers(8, 11, init.shape[:2]) // scale
    init[rng.random(init.shape[:2]) < 0.1, 0] = NOT_FOUND
    init[:8 // scale + 1, :40 // scale, 0] = 0
    got = orc.aw1_ref(L, R, sub=s, dmin=0, dmax=64, init=init, fill=0x4321)
    want = numpy_aw1_ref(L, R, s, 0, 0, 64, 0, True, 0.0, 0.0, 0.0, init=init, fill=0x4321)
    assert np.array_equal(got, want), int((got != want).any(-1).sum())
    assert (got[..., 0] == FIRST + 9).mean() > 0.3


def test_aw1_ref_known_shift_full_similarity():
    """R(x) = L(x + s): interior results find d = s with every pixel pair equal, so the
    similarity is 100 % (channel 1 = 65535)."""
    W, H, s = 160, 60, 11
    rng = np.random.default_rng(3)
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.zeros_like(L)
    R[:, :W - s] = L[:, s:]
    out = orc.aw1_ref(L, R, sub=5, dmin=0, dmax=32)
    core = out[20:40, 60:120]
    assert (core[..., 0] == FIRST + s).all()
    assert (core[..., 1] == 65535).all()


def test_aw1_ref_value_is_centre_plus_best_two_edge_neighbours():
    """d = 0 only: the similarity maximum is the centre sub-patch plus the two largest of
    the four edge neighbours, computed here by brute force for one result."""
    rng = np.random.default_rng(9)
    W, H, s = 80, 40, 3
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = rng.integers(0, 256, (H, W)).astype(np.uint8)
    out = orc.aw1_ref(L, R, sub=s, dmin=0, dmax=1)
    y, x = 12, 20  # a written centre: origin (8, 16) of block (0, 0) with 9 x 9 windows

    def S(cy, cx):
        a = L[cy - 1:cy + 2, cx - 1:cx + 2].astype(np.int64)
        b = R[cy - 1:cy + 2, cx - 1:cx + 2].astype(np.int64)
        return int((255 - np.abs(a - b)).sum())

    nb = sorted([S(y - s, x), S(y, x - s), S(y + s, x), S(y, x + s)])
    v = S(y, x) + nb[2] + nb[3]
    f = f32(v) / f32(27) * (f32(100.0) / f32(255.0))
    assert out[y, x, 1] == int(np.clip((f / f32(100) + f32(1)) / f32(2), 0, 1) * f32(65535.0))


def test_aw1_ref_written_area():
    """Blocks of 18 x 6 origins (15 x 15 windows), full blocks only."""
    W, H = 100, 50  # nbx = (100 - 14) // 18 = 4 (origins 0..71), nby = (50 - 14) // 6 = 6 (origins 0..35)
    L = _scene(7, W, H)[0]
    out = orc.aw1_ref(L, L, sub=5, dmin=0, dmax=16, fill=0x7777)
    ys, xs = np.nonzero((out != 0x7777).any(-1))
    assert xs.min() == 7 and ys.min() == 7
    assert xs.max() == 71 + 7 and ys.max() == min(35, H - 16) + 7
