"""The C++ driver layer (include/clutec/stereo.hpp) end to end: the stereo_run host
program configures and runs the drivers in the reference's call sequence; its maps
must equal the oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(cs.__file__), "lib", "stereo_run")


@pytest.mark.parametrize("algo,W,H,D", [("census", 320, 160, 64), ("sad", 300, 140, 32)])
def test_stereo_run_matches_oracle(tmp_path, algo, W, H, D):
    seed = 5
    prefix = str(tmp_path / algo)
    print(f"stereo_run {algo} {W}x{H} D={D}", flush=True)
    r = subprocess.run([EXE, algo, str(W), str(H), str(D), str(seed), prefix, "3"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    line = r.stdout.strip().splitlines()[-1].split()
    assert line[0] == "ms" and float(line[1]) > 0.0
    maps = [np.fromfile(prefix + s, np.int16).reshape(H, W) for s in ("_checked.i16", "_lr.i16", "_rl.i16")]
    L, R = cs.synth_pair(seed, W, H, 0, 8, D - 1, min(8, D - 1))
    if algo == "census":
        ref = orc.bm(orc.COST_CENSUS, orc.census(L, 9, 7), orc.census(R, 9, 7), D, window=(9, 9), census_win=(9, 7),
                     tol=1)
    else:
        ref = orc.bm(orc.COST_SAD, L, R, D, window=(9, 9), tol=1)
    for got, want in zip(maps, ref):
        assert np.array_equal(got, want)
    # the reference-style cross-check over the one-sided maps reports the kept range
    kept = maps[0][maps[0] >= 0]
    assert int(line[5]) == kept.min() + 1 and int(line[6]) == kept.max() + 1
    # DispDenoise (5x5 defaults) over the cross-check's id map
    ids = np.fromfile(prefix + "_checkid.u16x4", np.uint16).reshape(H, W, 4)
    den = np.fromfile(prefix + "_denoised.u16", np.uint16).reshape(H, W)
    assert np.array_equal(den, orc.disp_denoise(ids, P=5))
    # ColorDisparityImage (Analysis, ids 1 .. D + 1) and RectifyImage of the left image
    col = np.fromfile(prefix + "_color.rgba", np.uint8).reshape(H, W, 4)
    assert np.array_equal(col, orc.color_disparity(ids, 0, 1, D + 1))
    c, sn = np.float32(0.9998), np.float32(0.0199987)
    f, cx, cy = np.float32(0.8) * np.float32(W), np.float32(0.5) * np.float32(W - 1), np.float32(0.5) * np.float32(H - 1)
    p = orc.rectify_params((f, f, cx, cy), (-0.05, 0, 0, 0, 0), (f, f, cx, cy), [[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    rect = np.fromfile(prefix + "_rectified.u8", np.uint8).reshape(H, W)
    assert np.array_equal(rect, orc.rectify(L, p))


def test_stereo_run_bm1_matches_emulation(tmp_path):
    """BlockMatcher1::CDriver through the C++ layer (9x9, AbsoluteDifference, left to right)."""
    W, H, D, seed = 320, 160, 64, 7
    prefix = str(tmp_path / "bm1")
    print(f"stereo_run bm1 {W}x{H} D={D}", flush=True)
    r = subprocess.run([EXE, "bm1", str(W), str(H), str(D), str(seed), prefix, "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(prefix + "_bm1.u16x4", np.uint16).reshape(H, W, 4)
    L, R = cs.synth_pair(seed, W, H, 0, 8, D - 1, min(8, D - 1))
    want = orc.bm1(L, R, P=9, sim_type=0, dmin=0, dmax=D, offset=0, ltor=True, min_delta=1.0, sim_thresh=60.0,
                   grad_thresh=2.0, fill=0x5A5A)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("algo", ["aw1band", "nccband"])
def test_stereo_run_band_search_matches_oracle(tmp_path, algo):
    """Process(disp, L, R, dispInit) through the C++ layer with iUseDispInput (AW1 9x9 and
    NCC 9x9): the full-range map at half resolution as the initial map; AW1 bit-exact
    against the oracle restricted to the same per-tile bins, NCC equal but for FP32 near-ties."""
    from test_gpu_band import _tile_bins
    W, H, D, seed = 288, 150, 64, 7
    prefix = str(tmp_path / algo)
    print(f"stereo_run {algo} {W}x{H} D={D}", flush=True)
    r = subprocess.run([EXE, algo, str(W), str(H), str(D), str(seed), prefix, "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    init = np.fromfile(prefix + "_init.i16", np.int16).reshape(H // 2, W // 2)
    got = np.fromfile(prefix + "_band.i16", np.int16).reshape(H, W)
    L, R = cs.synth_pair(seed, W, H, 0, 8, D - 1, min(8, D - 1))
    bins = _tile_bins(init, W, H, D)
    if algo == "aw1band":
        assert np.array_equal(got, orc.aw_bins(1, L, R, D, bins, sub=3, tol=1, nthreads=8)[0])
    else:
        want = orc.ncc_bins(L, R, D, bins, window=(9, 9), tol=1, nthreads=8)[0]
        assert (got == want).mean() >= 0.999
    assert (got >= 0).mean() > 0.5


@pytest.mark.parametrize("algo", ["nccref", "aw1ref"])
def test_stereo_run_reference_operators_with_min_and_offset(tmp_path, algo):
    """BlockMatcherNCC / BlockMatcherAW1::CDriver through the C++ layer, configured with the
    reference's own SParameter::Set arguments over _SDisparityConfig [3, D) with pixel
    offset 2 and Process(disp, L, R, dispInit, debug): bit-exact against the emulations."""
    W, H, D, seed = 320, 160, 64, 9
    prefix = str(tmp_path / algo)
    print(f"stereo_run {algo} {W}x{H} D={D}", flush=True)
    r = subprocess.run([EXE, algo, str(W), str(H), str(D), str(seed), prefix, "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(prefix + "_ref.u16x4", np.uint16).reshape(H, W, 4)
    L, R = cs.synth_pair(seed, W, H, 0, 8, D - 1, min(8, D - 1))
    if algo == "nccref":
        want = orc.ncc_ref(L, R, P=9, dmin=3, dmax=D, offset=2, ltor=True, min_delta=2, ncc_thresh=0.5,
                           grad_thresh=2.0, fill=0x5A5A)
    else:
        want = orc.aw1_ref(L, R, sub=5, sim_type=0, dmin=3, dmax=D, offset=2, ltor=True, min_delta=0.5,
                           sim_thresh=60.0, grad_thresh=2.0, fill=0x5A5A)
    assert np.array_equal(got, want)
    ids = got[..., 0]
    assert ((ids >= 1) & (ids <= 0xFFEF)).mean() > 0.3
