"""The reference-format NCC and AW1 operators on the GPU (stereo_ncc_*, stereo_aw1_*) against
the oracle's block-by-block emulations (orc_ncc_ref, orc_aw1_ref), bit-exact in all four
channels, including the pixels the reference leaves unwritten; full frames of the NCC
(1920x1080, D=256) and adaptive-window (1280x720, D=128) workloads included."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc
from test_oracle_ref import _scene

pytestmark = pytest.mark.gpu

FILL = 0x5A5A
FIRST, LAST = 1, 0xFFEF


def _check(got, want, what):
    bad = (got != want).any(-1)
    assert not bad.any(), f"{what}: {int(bad.sum())} pixels differ, first at {tuple(np.argwhere(bad)[0])}"


def _ncc(ctx, L, R, init=None, **kw):
    H, W = L.shape
    m = cs.BlockMatcherNCC(ctx, W, H, use_disp_input=init is not None, **kw)
    try:
        return m.run(L, R, init=init, fill=FILL)
    finally:
        m.close()


def _ncc_oracle(L, R, init=None, patch=11, disp_min=0, disp_max=64, disp_offset=0, min_delta_thresh=0,
                ncc_thresh=0.0, grad_thresh=0.0, left_to_right=True):
    return orc.ncc_ref(L, R, P=patch, dmin=disp_min, dmax=disp_max, offset=disp_offset, ltor=left_to_right,
                       min_delta=min_delta_thresh, ncc_thresh=ncc_thresh, grad_thresh=grad_thresh, init=init,
                       fill=FILL)


def _aw1(ctx, L, R, init=None, **kw):
    H, W = L.shape
    m = cs.BlockMatcherAW1(ctx, W, H, use_disp_input=init is not None, **kw)
    try:
        return m.run(L, R, init=init, fill=FILL)
    finally:
        m.close()


def _aw1_oracle(L, R, init=None, patch=15, sim_type=0, disp_min=0, disp_max=64, disp_offset=0,
                min_delta_thresh=0.0, sim_thresh=0.0, grad_thresh=0.0, left_to_right=True):
    return orc.aw1_ref(L, R, sub=patch // 3, sim_type=sim_type, dmin=disp_min, dmax=disp_max, offset=disp_offset,
                       ltor=left_to_right, min_delta=min_delta_thresh, sim_thresh=sim_thresh,
                       grad_thresh=grad_thresh, init=init, fill=FILL)


NCC_CASES = [
    # patch, ltor, dmin, dmax, offset, (iMinDeltaThresh, fNccThresh, fGradThresh)
    (11, True, 0, 64, 0, (2, 0.6, 3.0)),
    (9, True, 2, 40, 1, (3, 0.5, 4.0)),
    (5, True, 0, 96, 0, (1, 0.7, 2.0)),
    (11, False, 2, 40, 1, (2, 0.6, 3.0)),
    (9, False, 5, 70, -2, (4, 0.8, 5.0)),
    (5, True, 3, 51, 3, (0, 0.0, 0.0)),
]


@pytest.mark.parametrize("patch,ltor,dmin,dmax,off,thr", NCC_CASES)
def test_ncc_ref_matches_emulation(ctx, patch, ltor, dmin, dmax, off, thr):
    W, H = 331, 203
    print(f"ncc_ref P{patch} ltor{int(ltor)} [{dmin},{dmax}) off {off}", flush=True)
    L, R = _scene(patch, W, H, shifts=(9, 27), ltor=ltor)
    kw = dict(patch=patch, disp_min=dmin, disp_max=dmax, disp_offset=off, min_delta_thresh=thr[0],
              ncc_thresh=thr[1], grad_thresh=thr[2], left_to_right=ltor)
    got = _ncc(ctx, L, R, **kw)
    want = _ncc_oracle(L, R, **kw)
    _check(got, want, "map")
    c = want[..., 0]
    assert ((c >= FIRST) & (c <= LAST)).mean() > 0.2


@pytest.mark.parametrize("W,H,patch", [(48, 20, 5), (49, 26, 11), (70, 40, 9), (131, 47, 11)])
def test_ncc_ref_small_and_ragged(ctx, W, H, patch):
    print(f"ncc_ref small {W}x{H} P{patch}", flush=True)
    L, R = _scene(W + patch, W, H, shifts=(3, 6))
    kw = dict(patch=patch, disp_min=0, disp_max=32, min_delta_thresh=1, ncc_thresh=0.3, grad_thresh=1.0)
    _check(_ncc(ctx, L, R, **kw), _ncc_oracle(L, R, **kw), f"{W}x{H}")


@pytest.mark.parametrize("patch,scale", [(11, 2), (9, 4), (5, 1)])
def test_ncc_ref_band_search_matches_emulation(ctx, patch, scale):
    W, H = 300, 160
    print(f"ncc_ref band P{patch} scale {scale}", flush=True)
    L = _scene(50 + patch, W, H)[0]
    R = np.roll(L, -19, axis=1)
    rng = np.random.default_rng(patch)
    init = np.zeros((H // scale, W // scale, 4), np.uint16)
    init[..., 0] = FIRST + rng.integers(17, 22, init.shape[:2]) // scale
    init[rng.random(init.shape[:2]) < 0.1, 0] = 0xFFFD
    init[:8 // scale + 1, :40 // scale, 0] = 0
    kw = dict(patch=patch, disp_min=0, disp_max=64, min_delta_thresh=1, ncc_thresh=0.3)
    got = _ncc(ctx, L, R, init=init, **kw)
    want = _ncc_oracle(L, R, init=init, **kw)
    _check(got, want, "band")
    assert (want[..., 0] == FIRST + 19).mean() > 0.3


def test_ncc_ref_full_frame_1080p_d256(ctx):
    """The NCC workload's frame (1920x1080, D=256, 11x11) through the reference operator:
    bit-exact against the emulation, and most written ids equal the plain fused matcher's
    left map (which takes the exact score's argmax over the full precision)."""
    W, H, D = 1920, 1080, 256
    print("ncc_ref 1080p D=256", flush=True)
    L, R = cs.synth_pair(1, W, H, 0, 24, D - 1, 8)
    kw = dict(patch=11, disp_min=0, disp_max=D, min_delta_thresh=0, ncc_thresh=0.0, grad_thresh=0.0)
    got = _ncc(ctx, L, R, **kw)
    want = orc.ncc_ref(L, R, P=11, dmin=0, dmax=D, fill=FILL)
    _check(got, want, "1080p")
    ids = got[..., 0].astype(np.int64)
    plain = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_NCC, window=(11, 11), num_disp=D, tol=1)
    _, lr, _ = plain.run(L, R, maps=True)
    valid = (ids >= FIRST) & (ids <= LAST) & (lr >= 0)
    agree = float((ids[valid] - FIRST == lr[valid]).mean())
    print(f"reference-format NCC vs plain fused NCC (left maps): {agree:.4f} of {int(valid.sum())} ids", flush=True)
    assert valid.mean() > 0.5 and agree > 0.9


AW_CASES = [
    # patch, sim, ltor, dmin, dmax, offset, (min_delta, sim_thresh, grad_thresh)
    (15, 0, True, 0, 64, 0, (1.0, 80.0, 3.0)),
    (9, 0, True, 2, 40, 1, (0.5, 75.0, 2.0)),
    (15, 1, True, 0, 48, 0, (0.3, 5.0, 3.0)),
    (9, 1, False, 2, 40, 1, (0.5, 5.0, 2.0)),
    (15, 0, False, 5, 70, -2, (1.0, 80.0, 4.0)),
    (9, 0, True, 0, 96, 0, (0.0, 0.0, 0.0)),
]


@pytest.mark.parametrize("patch,sim,ltor,dmin,dmax,off,thr", AW_CASES)
def test_aw1_ref_matches_emulation(ctx, patch, sim, ltor, dmin, dmax, off, thr):
    W, H = 331, 203
    print(f"aw1_ref P{patch} sim{sim} ltor{int(ltor)} [{dmin},{dmax}) off {off}", flush=True)
    L, R = _scene(60 + patch, W, H, shifts=(9, 27), ltor=ltor)
    kw = dict(patch=patch, sim_type=sim, disp_min=dmin, disp_max=dmax, disp_offset=off, min_delta_thresh=thr[0],
              sim_thresh=thr[1], grad_thresh=thr[2], left_to_right=ltor)
    got = _aw1(ctx, L, R, **kw)
    want = _aw1_oracle(L, R, **kw)
    _check(got, want, "map")
    c = want[..., 0]
    assert ((c >= FIRST) & (c <= LAST)).mean() > 0.2


@pytest.mark.parametrize("W,H,patch", [(48, 20, 15), (49, 18, 9), (97, 33, 15), (130, 64, 9)])
def test_aw1_ref_small_and_ragged(ctx, W, H, patch):
    print(f"aw1_ref small {W}x{H} P{patch}", flush=True)
    L, R = _scene(W + patch, W, H, shifts=(3, 6))
    kw = dict(patch=patch, disp_min=0, disp_max=32, min_delta_thresh=0.5, sim_thresh=60.0, grad_thresh=1.0)
    _check(_aw1(ctx, L, R, **kw), _aw1_oracle(L, R, **kw), f"{W}x{H}")


@pytest.mark.parametrize("patch,scale", [(15, 2), (9, 4)])
def test_aw1_ref_band_search_matches_emulation(ctx, patch, scale):
    W, H = 300, 160
    print(f"aw1_ref band P{patch} scale {scale}", flush=True)
    L = _scene(70 + patch, W, H)[0]
    R = np.roll(L, -19, axis=1)
    rng = np.random.default_rng(patch)
    init = np.zeros((H // scale, W // scale, 4), np.uint16)
    init[..., 0] = FIRST + rng.integers(17, 22, init.shape[:2]) // scale
    init[rng.random(init.shape[:2]) < 0.1, 0] = 0xFFFD
    init[:8 // scale + 1, :40 // scale, 0] = 0
    kw = dict(patch=patch, disp_min=0, disp_max=64)
    got = _aw1(ctx, L, R, init=init, **kw)
    want = _aw1_oracle(L, R, init=init, **kw)
    _check(got, want, "band")
    assert (want[..., 0] == FIRST + 19).mean() > 0.3


def test_aw1_ref_full_frame_720p_d128(ctx):
    """The adaptive-window workload's frame (1280x720, slanted planes, D=128, 15x15) through the
    reference operator, bit-exact against the emulation."""
    W, H, D = 1280, 720, 128
    print("aw1_ref 720p D=128", flush=True)
    L, R = cs.synth_pair(2, W, H, 1, 6, D - 1, 8)
    kw = dict(patch=15, disp_min=0, disp_max=D, min_delta_thresh=0.5, sim_thresh=60.0, grad_thresh=2.0)
    got = _aw1(ctx, L, R, **kw)
    want = _aw1_oracle(L, R, **kw)
    _check(got, want, "720p")
    c = got[..., 0]
    assert ((c >= FIRST) & (c <= LAST)).mean() > 0.5
