"""BlockMatcher1 on the GPU against the oracle's block-by-block emulation of the reference,
bit-exact in all four channels, including the pixels the reference leaves unwritten."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc
from test_oracle_bm1 import _scene

pytestmark = pytest.mark.gpu

FILL = 0x5A5A


def _run(ctx, L, R, init=None, **kw):
    H, W = L.shape
    m = cs.BlockMatcher1(ctx, W, H, use_disp_input=init is not None, **kw)
    return m.run(L, R, init=init, fill=FILL)


def _oracle(L, R, init=None, patch=9, sim_type=0, disp_min=0, disp_max=64, disp_offset=0, min_delta_thresh=0.0,
            sim_thresh=0.0, grad_thresh=0.0, left_to_right=True):
    return orc.bm1(L, R, P=patch, sim_type=sim_type, dmin=disp_min, dmax=disp_max, offset=disp_offset,
                   ltor=left_to_right, min_delta=min_delta_thresh, sim_thresh=sim_thresh, grad_thresh=grad_thresh,
                   init=init, fill=FILL)


def _check(got, want, what):
    bad = (got != want).any(-1)
    assert not bad.any(), f"{what}: {int(bad.sum())} pixels differ, first at {tuple(np.argwhere(bad)[0])}"


CASES = [
    # patch, sim, ltor, dmin, dmax, offset, (min_delta, sim_thresh, grad_thresh)
    (5, 0, True, 0, 64, 0, (2.0, 70.0, 4.0)),
    (9, 0, True, 2, 40, 1, (2.0, 70.0, 4.0)),
    (11, 0, True, 0, 96, 0, (1.0, 60.0, 3.0)),
    (15, 0, True, 5, 70, 2, (2.0, 70.0, 4.0)),
    (5, 1, True, 0, 48, 0, (0.5, 5.0, 4.0)),
    (9, 1, False, 2, 40, 1, (0.5, 5.0, 4.0)),
    (11, 1, True, 0, 64, -1, (0.5, 5.0, 4.0)),
    (15, 0, False, 0, 64, 0, (2.0, 70.0, 4.0)),
]


@pytest.mark.parametrize("patch,sim,ltor,dmin,dmax,off,thr", CASES)
def test_bm1_matches_emulation(ctx, patch, sim, ltor, dmin, dmax, off, thr):
    W, H = 331, 203
    print(f"bm1 P{patch} sim{sim} ltor{int(ltor)} [{dmin},{dmax}) off {off}", flush=True)
    L, R = _scene(patch, W, H, shifts=(9, 27), ltor=ltor)
    kw = dict(patch=patch, sim_type=sim, disp_min=dmin, disp_max=dmax, disp_offset=off, min_delta_thresh=thr[0],
              sim_thresh=thr[1], grad_thresh=thr[2], left_to_right=ltor)
    got = _run(ctx, L, R, **kw)
    want = _oracle(L, R, **kw)
    _check(got, want, "map")
    c = want[..., 0]
    assert ((c >= 1) & (c <= 0xFFEF)).mean() > 0.3  # a real share of results carries a disparity


@pytest.mark.parametrize("W,H,patch", [(48, 8, 5), (49, 12, 9), (71, 14, 11), (97, 18, 15), (130, 64, 9)])
def test_bm1_edge_sizes(ctx, W, H, patch):
    print(f"bm1 edge {W}x{H} P{patch}", flush=True)
    L, R = _scene(W + H, W, H, shifts=(3, 7))
    kw = dict(patch=patch, disp_min=0, disp_max=32, grad_thresh=2.0, sim_thresh=50.0, min_delta_thresh=1.0)
    _check(_run(ctx, L, R, **kw), _oracle(L, R, **kw), "edge")


def test_bm1_640x480(ctx):
    W, H = 640, 480
    print("bm1 640x480 P9 D64", flush=True)
    L, R = cs.synth_pair(0, W, H, 0, 8, 63, 8)
    kw = dict(patch=9, disp_min=0, disp_max=64, min_delta_thresh=1.0, sim_thresh=60.0, grad_thresh=2.0)
    got = _run(ctx, L, R, **kw)
    _check(got, _oracle(L, R, **kw), "640x480")
    # the synthetic pair's disparities are found on most of the textured frame
    c = got[..., 0]
    assert ((c >= 1) & (c <= 64)).mean() > 0.5


@pytest.mark.parametrize("scale,patch", [(2, 9), (1, 5), (4, 11)])
def test_bm1_band_search_matches_emulation(ctx, scale, patch):
    W, H = 256, 120
    print(f"bm1 band search scale {scale} P{patch}", flush=True)
    L, R = _scene(40 + scale, W, H, shifts=(6, 14))
    rng = np.random.default_rng(scale)
    ih, iw = H // scale, W // scale
    init = np.zeros((ih, iw, 4), np.uint16)
    init[..., 0] = 1 + rng.integers(1, 14, (ih, iw)) // scale
    init[rng.random((ih, iw)) < 0.1, 0] = 0xFFFD
    init[: 8 // scale + 1, : 40 // scale, 0] = 0  # blocks whose threads see Unknown
    kw = dict(patch=patch, disp_min=0, disp_max=64, min_delta_thresh=1.0, sim_thresh=50.0)
    _check(_run(ctx, L, R, init=init, **kw), _oracle(L, R, init=init, **kw), "band")


def test_bm1_configuration_errors(ctx):
    with pytest.raises(cs.StereoError) as e:
        cs.BlockMatcher1(ctx, 128, 64, patch=7)
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    with pytest.raises(cs.StereoError) as e:
        cs.BlockMatcher1(ctx, 40, 64, patch=5)
    assert e.value.status == cs.ERROR_INVALID_FORMAT
    with pytest.raises(cs.StereoError) as e:
        cs.BlockMatcher1(ctx, 128, 64, disp_min=10, disp_max=10)
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    m = cs.BlockMatcher1(ctx, 128, 64, patch=9, use_disp_input=True)
    L = np.zeros((64, 128), np.uint8)
    with pytest.raises(cs.StereoError):
        m.run(L, L)  # no initial map
