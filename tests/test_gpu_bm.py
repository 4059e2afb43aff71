"""GPU parity of the census transform, the fused SAD / census matchers and the
reference-format cross check against the CPU oracle: bit-exact on every map."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu


def _pair(seed, W, H, kind=0, nreg=8, dmax=63):
    return cs.synth_pair(seed, W, H, kind, nreg, dmax, min(8, dmax))


# ------------------------------------------------------------------- census

@pytest.mark.parametrize("win", [(9, 7), (7, 9), (5, 5), (7, 7), (3, 3)])
@pytest.mark.parametrize("size", [(131, 77), (640, 480)])
def test_census_wide(ctx, win, size):
    W, H = size
    L, _ = _pair(11, W, H)
    got = cs.Census(ctx, W, H, window=win).run(L)
    assert np.array_equal(got, orc.census(L, *win))


def test_census_reference_3x3(ctx):
    L, _ = _pair(12, 203, 61)
    got = cs.Census(ctx, 203, 61, window=(3, 3), reference_3x3=True).run(L)
    assert np.array_equal(got, orc.census3x3_ref(L))


def test_census_format_errors(ctx):
    c = cs.Census(ctx, 64, 32)
    src = cs.DeviceImage(ctx, 64, 32, cs.PIXEL_LUM_U8)
    wrong_type = cs.DeviceImage(ctx, 64, 32, cs.PIXEL_LUM_U32)
    with pytest.raises(cs.StereoError) as e:
        c.process(wrong_type, src)
    assert e.value.status == cs.ERROR_UNSUPPORTED_PIXEL_TYPE
    wrong_size = cs.DeviceImage(ctx, 60, 32, cs.PIXEL_LUM_U64)
    with pytest.raises(cs.StereoError) as e:
        c.process(wrong_size, src)
    assert e.value.status == cs.ERROR_FORMAT_MISMATCH


# ------------------------------------------------------------------- matchers

def _oracle(algo, L, R, D, win, cen, tol):
    if algo == cs.ALGO_SAD:
        return orc.bm(orc.COST_SAD, L, R, D, window=win, tol=tol, nthreads=8)
    return orc.bm(orc.COST_CENSUS, orc.census(L, *cen), orc.census(R, *cen), D, window=win, census_win=cen,
                  tol=tol, nthreads=8)


CASES = [
    # algo, W, H, D, window, census
    (cs.ALGO_SAD, 200, 120, 64, (9, 9), (1, 1)),
    (cs.ALGO_SAD, 333, 97, 128, (7, 7), (1, 1)),
    (cs.ALGO_SAD, 150, 70, 40, (5, 5), (1, 1)),
    (cs.ALGO_CENSUS, 200, 120, 64, (9, 9), (9, 7)),
    (cs.ALGO_CENSUS, 419, 133, 192, (9, 9), (9, 7)),
    (cs.ALGO_CENSUS, 300, 90, 256, (7, 7), (7, 7)),
    (cs.ALGO_CENSUS, 130, 60, 16, (5, 5), (5, 5)),
    # every census window the matchers take: the matchers' descriptors pair neighbours
    # differently for each width (odd widths pair the last column across rows)
    (cs.ALGO_CENSUS, 250, 110, 96, (9, 9), (7, 9)),
    (cs.ALGO_CENSUS, 170, 75, 48, (5, 5), (3, 3)),
]


@pytest.mark.parametrize("algo,W,H,D,win,cen", CASES)
def test_matcher_bit_exact(ctx, algo, W, H, D, win, cen):
    print(f"matcher algo={algo} {W}x{H} D={D} win={win} census={cen}", flush=True)
    L, R = _pair(W * 7 + H, W, H, 0, 6, max(8, D - 1))
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=win, census=cen, num_disp=D, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ref_out, ref_l, ref_r = _oracle(algo, L, R, D, win, cen, 1)
    assert np.array_equal(lr, ref_l)
    assert np.array_equal(rl, ref_r)
    assert np.array_equal(out, ref_out)


def test_sad_config0_full(ctx):
    """640x480, D=64, 9x9 SAD, tolerance 1 (the CPU-runnable workload)."""
    L, R = _pair(0, 640, 480, 0, 8, 63)
    m = cs.Matcher(ctx, 640, 480, algorithm=cs.ALGO_SAD, window=(9, 9), num_disp=64, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ref = _oracle(cs.ALGO_SAD, L, R, 64, (9, 9), (1, 1), 1)
    assert np.array_equal(lr, ref[1]) and np.array_equal(rl, ref[2]) and np.array_equal(out, ref[0])


def test_census_config1_full(ctx):
    """1920x1080, D=256, census 9x7, 9x9 aggregation (the headline workload)."""
    L, R = cs.synth_pair(1, 1920, 1080, 0, 24, 255, 8)
    m = cs.Matcher(ctx, 1920, 1080, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=256, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ref = _oracle(cs.ALGO_CENSUS, L, R, 256, (9, 9), (9, 7), 1)
    assert np.array_equal(lr, ref[1])
    assert np.array_equal(rl, ref[2])
    assert np.array_equal(out, ref[0])
    # sanity against the generator's ground truth: most checked pixels are right
    _, _, gt = cs.synth_pair(1, 1920, 1080, 0, 24, 255, 8, with_gt=True)
    v = out >= 0
    assert v.mean() > 0.5 and (out[v] == gt[v]).mean() > 0.9


def test_matcher_disparity_id_output(ctx):
    W, H, D = 180, 90, 48
    L, R = _pair(21, W, H, 0, 5, 47)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1)
    ids = m.run(L, R, dispid=True)
    out, lr, rl, cost = orc.bm(orc.COST_CENSUS, orc.census(L), orc.census(R), D, census_win=(9, 7), tol=1,
                               with_cost=True)
    border = np.ones((H, W), bool)
    border[7:H - 7, 8:W - 8] = False
    expect = np.where(border, orc.DISP_UNKNOWN,
                      np.where(out >= 0, out + 1, np.where(lr >= 0, orc.DISP_INCONSISTENT, orc.DISP_NOT_FOUND)))
    assert np.array_equal(ids[:, :, 0], expect.astype(np.uint16))
    assert np.array_equal(ids[:, :, 1][lr >= 0], cost[lr >= 0].astype(np.uint16))


def test_matcher_errors(ctx):
    W, H = 128, 64
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_SAD, window=(9, 9), num_disp=32)
    left = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U8)
    right_bad = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U16)
    out = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
    with pytest.raises(cs.StereoError) as e:
        m.process(out, None, None, left, right_bad)
    assert e.value.status == cs.ERROR_UNSUPPORTED_PIXEL_TYPE
    small = cs.DeviceImage(ctx, W - 1, H, cs.PIXEL_LUM_U8)
    with pytest.raises(cs.StereoError) as e:
        m.process(out, None, None, left, small)
    assert e.value.status == cs.ERROR_FORMAT_MISMATCH
    with pytest.raises(cs.StereoError) as e:
        cs.Matcher(ctx, W, H, algorithm=7)
    assert e.value.status == cs.ERROR_INVALID_CONFIG


def test_matcher_timing(ctx):
    W, H = 256, 128
    L, R = _pair(5, W, H)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, num_disp=64)
    ctx.set_timing(True)
    try:
        m.run(L, R)
        assert m.last_time_ms() > 0.0
    finally:
        ctx.set_timing(False)


# ------------------------------------------------------------------- cross check

def _random_id_map(rng, W, H, D):
    m = np.zeros((H, W, 4), np.uint16)
    ids = rng.integers(1, D + 1, (H, W))
    special = rng.random((H, W))
    ids = np.where(special < 0.05, orc.DISP_NOT_FOUND, ids)
    ids = np.where((special > 0.05) & (special < 0.08), orc.DISP_UNKNOWN, ids)
    m[:, :, 0] = ids
    m[:, :, 1:] = rng.integers(0, 65536, (H, W, 3))
    return m


@pytest.mark.parametrize("ltor,thresh,offset", [(1, 1, 0), (0, 0, 0), (1, 2, 3)])
def test_crosscheck_reference(ctx, ltor, thresh, offset):
    rng = np.random.default_rng(ltor * 10 + thresh)
    W, H, D = 157, 43, 12
    lr = _random_id_map(rng, W, H, D)
    rl = lr.copy()
    rl[:, :, 0] = np.clip(lr[:, :, 0].astype(int) + rng.integers(-2, 3, (H, W)), 0, 65535)
    cc = cs.CrossCheck(ctx, W, H, thresh=thresh, left_to_right=bool(ltor), offset=offset)
    got = cc.run(lr, rl)
    ref, mn, mx = orc.crosscheck_ref(lr, rl, thresh, ltor, offset)
    assert np.array_equal(got, ref)
    assert cc.range() == (mn, mx)


def test_crosscheck_format_mismatch(ctx):
    cc = cs.CrossCheck(ctx, 32, 16)
    a = cs.DeviceImage(ctx, 32, 16, cs.PIXEL_RGBA_U16)
    b = cs.DeviceImage(ctx, 31, 16, cs.PIXEL_RGBA_U16)
    with pytest.raises(cs.StereoError) as e:
        cc.process(a, a, b)
    assert e.value.status == cs.ERROR_FORMAT_MISMATCH


@pytest.mark.parametrize("op", [0, 1])
def test_valu_peak_probe(ctx, op):
    """The measured INT32 issue rate is positive and within a factor of the
    specified 39.3 T lane-ops/s (256 CUs x 64 lanes x 2.4 GHz)."""
    v = ctx.valu_peak(op)
    print(f"valu peak op {op}: {v:.2f} T lane-ops/s", flush=True)
    assert 5.0 < v < 100.0
