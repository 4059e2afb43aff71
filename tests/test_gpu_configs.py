"""The benchmark workloads at their full sizes on the GPU against the oracle:
  - the batch of 64 pairs of uneven size and range (W = 16 U[20, 128], H = 16 U[15, 96],
    D in {64, 128, 192, 256}, seeds 100-163, the same pairs as the benchmark), SAD 7x7 and
    census 9x7 / 9x9, every pair's three maps bit-exact;
  - the learned proposal at 1920x1080, D=256, with the trained network: maps equal to the
    oracle restricted to the searched bins, agreement with the full range >= 0.995, and the
    network's gain over the same method without it;
  - the fused NCC (11x11) on the full 1920x1080 D=256 frame and the fused AW1 on the full
    1280x720 D=128 slanted-plane frame."""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)))
WEIGHTS = os.path.join(os.path.dirname(cs.__file__), "weights", "proposal_cnn.bin")


def _batch_pairs():
    rng = np.random.default_rng(100)
    sizes = [(16 * int(rng.integers(20, 129)), 16 * int(rng.integers(15, 97))) for _ in range(64)]
    disps = [int(rng.choice([64, 128, 192, 256])) for _ in range(64)]
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        pairs = list(ex.map(lambda i: cs.synth_pair(100 + i, sizes[i][0], sizes[i][1], 0, 8, disps[i] - 1, 8),
                            range(64)))
    return sizes, disps, pairs


@pytest.mark.parametrize("name", ["sad7x7", "census9x7"])
def test_batch_64_pairs_match_oracle(ctx, name):
    sizes, disps, pairs = _batch_pairs()
    print(f"batch {name}: 64 pairs, {sum(w * h * d for (w, h), d in zip(sizes, disps)) / 1e9:.2f} G evals", flush=True)
    if name == "sad7x7":
        b = cs.BatchMatcher(ctx, sizes, disps, algorithm=cs.ALGO_SAD, window=(7, 7))
    else:
        b = cs.BatchMatcher(ctx, sizes, disps, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7))
    got = b.run(pairs)
    b.process()  # the second call replays the captured graph
    again = b.results()
    for i, ((L, R), d) in enumerate(zip(pairs, disps)):
        if name == "sad7x7":
            ref = orc.bm(orc.COST_SAD, L, R, d, window=(7, 7), tol=1, nthreads=THREADS)
        else:
            ref = orc.bm(orc.COST_CENSUS, orc.census(L, 9, 7), orc.census(R, 9, 7), d, window=(9, 9),
                         census_win=(9, 7), tol=1, nthreads=THREADS)
        for k, what in enumerate(("checked", "left", "right")):
            assert np.array_equal(got[i][k], ref[k]), f"pair {i} {sizes[i]} D={d}: {what} map"
            assert np.array_equal(again[i][k], ref[k]), f"pair {i} replay: {what} map"
    b.close()


def _learned_params():
    params = np.fromfile(WEIGHTS, "<f4")
    assert params.size == cs.CNN_PARAMS
    with open(WEIGHTS + ".json") as f:
        meta = json.load(f)
    return params, meta


def test_learned_1080p_trained_network(ctx):
    """The learned-proposal matcher on the 1080p D=256 frame with the committed network:
    exact on the bins it searched, >= 99.5 % equal to the full range, and better than the
    method without the network -- the same coarse bands with no tile sent back to the full
    range, and the reference's own rule (bins of [min - 4, max + 4] around the quarter-
    resolution census map, MatcherUseDisparity)."""
    W, H, D = 1920, 1080, 256
    L, R = cs.synth_pair(1, W, H, 0, 24, D - 1, 8)
    params, meta = _learned_params()
    kw = dict(num_disp=D, band=int(meta.get("band", 8)), chunk=int(meta.get("chunk", 32)),
              centre_residual=bool(meta.get("centre_residual", 0)))
    thr = float(meta["logit_thresh"])
    print(f"learned 1080p D=256, threshold {thr}", flush=True)
    lm = cs.LearnedMatcher(ctx, W, H, params, logit_thresh=thr, **kw)
    got = lm.run(L, R, maps=True)
    st = lm.stats()
    masks = lm.tile_masks()
    bins = np.repeat(np.repeat(masks, 56, axis=0), 64, axis=1)[:H, :W]
    cl, cr = orc.census(L, 9, 7), orc.census(R, 9, 7)
    ref = orc.bm_bins(orc.COST_CENSUS, cl, cr, D, bins, window=(9, 9), census_win=(9, 7), tol=1, nthreads=THREADS)
    for a, c, what in zip(got, ref, ("checked", "left", "right")):
        assert np.array_equal(a, c), f"learned {what} map differs from the oracle on its bins"
    full = orc.bm(orc.COST_CENSUS, cl, cr, D, window=(9, 9), census_win=(9, 7), tol=1, nthreads=THREADS)[0]
    agree = float((got[0] == full).mean())

    nonet = cs.LearnedMatcher(ctx, W, H, params, logit_thresh=-1e30, **kw)
    agree_nonet = float((nonet.run(L, R) == full).mean())
    st_nonet = nonet.stats()

    def ds4(img):
        return ((img.reshape(H // 4, 4, W // 4, 4).astype(np.int32).sum(axis=(1, 3)) + 8) >> 4).astype(np.uint8)

    coarse = cs.Matcher(ctx, W // 4, H // 4, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7),
                        num_disp=D // 4, tol=1)
    dc = coarse.run(ds4(L), ds4(R))
    band = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1,
                      use_disp_input=True)
    agree_rule = float((band.run_init(L, R, dc) == full).mean())
    print(f"agreement with the full range: learned {agree:.5f} ({st['fallback_tiles']}/{st['tiles']} tiles back to "
          f"the full range, {st['evals'] / (W * H * D):.3f} of the evaluations); no network {agree_nonet:.5f} "
          f"({st_nonet['evals'] / (W * H * D):.3f}); reference rule {agree_rule:.5f}", flush=True)
    assert agree >= 0.995
    assert agree > agree_nonet and agree > agree_rule
    assert st["evals"] < 0.5 * W * H * D


def test_ncc_full_frame_1080p_d256(ctx):
    """The fused NCC on the whole 1920x1080 frame, D=256, 11x11: every winning score within
    1e-5 of FP64, the maps equal but for FP32 near-ties (pixels counted)."""
    W, H, D = 1920, 1080, 256
    L, R = cs.synth_pair(1, W, H, 0, 24, D - 1, 8)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_NCC, window=(11, 11), num_disp=D, tol=1)
    out, lr, rl, score = m.run(L, R, maps=True, scores=True)
    ref_out, ref_l, ref_r, ref_sc = orc.ncc(L, R, D, window=(11, 11), tol=1, nthreads=THREADS)
    same = (lr == ref_l) & (lr >= 0)
    nz = same & (ref_sc != 0)
    rel = np.abs(score[nz] - ref_sc[nz]) / np.abs(ref_sc[nz])
    print(f"NCC 1080p: left {(lr == ref_l).mean():.6f}, right {(rl == ref_r).mean():.6f}, checked "
          f"{(out == ref_out).mean():.6f} equal; {int((lr != ref_l).sum())} left pixels differ; max rel score "
          f"error {rel.max():.3e}", flush=True)
    assert rel.max() <= 1e-5
    assert np.array_equal(lr >= 0, ref_l >= 0)
    assert (lr == ref_l).mean() >= 0.999 and (rl == ref_r).mean() >= 0.999 and (out == ref_out).mean() >= 0.998


def test_aw1_full_frame_720p_d128(ctx):
    """The fused AW1 (similarity 255 - |a - b|, 5x5 sub-patches) on the whole slanted-plane
    1280x720 frame, D=128: bit-exact."""
    L, R = cs.synth_pair(2, 1280, 720, 1, 6, 127, 8)
    m = cs.Matcher(ctx, 1280, 720, algorithm=cs.ALGO_AW1, window=(5, 5), num_disp=128, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ref_out, ref_l, ref_r, _ = orc.aw(1, L, R, 128, sub=5, tol=1, nthreads=THREADS)
    assert np.array_equal(lr, ref_l) and np.array_equal(rl, ref_r) and np.array_equal(out, ref_out)
