"""Key planes across frames.  The census / SAD matchers fold every frame's keys into two
planes with atomicMax, and each frame's merge reads them and clears them for the next
frame.  Here two different pairs alternate on one driver, so a merge that left keys
behind, or read another frame's, would change a map: every frame must equal its own
pair's result.  The fused matcher, the band search and the batch each keep their own
planes."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc
from test_gpu_band import _tile_bins

pytestmark = pytest.mark.gpu

NFRAMES = 24


def _want(algo, L, R, D, window, census, bins=None):
    if algo == cs.ALGO_CENSUS:
        cl, cr = orc.census(L, *census), orc.census(R, *census)
        if bins is not None:
            return orc.bm_bins(orc.COST_CENSUS, cl, cr, D, bins, window=window, census_win=census, tol=1, nthreads=8)
        return orc.bm(orc.COST_CENSUS, cl, cr, D, window=window, census_win=census, tol=1, nthreads=8)
    if bins is not None:
        return orc.bm_bins(orc.COST_SAD, L, R, D, bins, window=window, tol=1, nthreads=8)
    return orc.bm(orc.COST_SAD, L, R, D, window=window, tol=1, nthreads=8)


def _frames(ctx, W, H, run_one, want):
    """Runs NFRAMES frames, pair f % 2 in frame f, and checks every map of every frame."""
    out, lr, rl = (cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16) for _ in range(3))
    for f in range(NFRAMES):
        k = f % 2
        run_one(k, out, lr, rl)
        got = (out.download(), lr.download(), rl.download())
        for a, b in zip(got, want[k]):
            assert np.array_equal(a, b), f"frame {f} (pair {k})"


@pytest.mark.parametrize("algo,window,census", [(cs.ALGO_CENSUS, (9, 9), (9, 7)), (cs.ALGO_SAD, (7, 7), (1, 1))])
def test_fused_matcher_planes_across_frames(ctx, algo, window, census):
    W, H, D = 160, 96, 32
    print(f"fused matcher planes algo={algo}", flush=True)
    pairs = [cs.synth_pair(s, W, H, 0, 6, D - 1, 8) for s in (5, 6)]
    want = [_want(algo, L, R, D, window, census) for L, R in pairs]
    assert not np.array_equal(want[0][0], want[1][0])
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=window, census=census, num_disp=D, tol=1)
    dev = [(cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8), cs.DeviceImage.from_array(ctx, R, cs.PIXEL_LUM_U8))
           for L, R in pairs]
    _frames(ctx, W, H, lambda k, out, lr, rl: m.process(out, lr, rl, *dev[k]), want)


def test_band_search_planes_across_frames(ctx):
    W, H, D = 192, 112, 64
    print("band search planes", flush=True)
    pairs = [cs.synth_pair(s, W, H, 2, 6, D - 1, 8) for s in (7, 8)]
    inits = []
    for L, R in pairs:
        full = _want(cs.ALGO_CENSUS, L, R, D, (9, 9), (9, 7))[0]
        inits.append((full[::2, ::2][: H // 2, : W // 2] // 2).astype(np.int16))
    want = [_want(cs.ALGO_CENSUS, L, R, D, (9, 9), (9, 7), _tile_bins(init, W, H, D))
            for (L, R), init in zip(pairs, inits)]
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1,
                   use_disp_input=True)
    dev = [(cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8), cs.DeviceImage.from_array(ctx, R, cs.PIXEL_LUM_U8),
            cs.DeviceImage.from_array(ctx, init, cs.PIXEL_LUM_I16)) for (L, R), init in zip(pairs, inits)]
    _frames(ctx, W, H, lambda k, out, lr, rl: m.process_init(out, lr, rl, *dev[k]), want)


def test_batch_planes_across_frames(ctx):
    """Two pairs per batch; the batch's content alternates between two sets of pairs."""
    sizes, disps = [(144, 80), (96, 64)], [32, 48]
    print("batch planes", flush=True)
    sets = [[cs.synth_pair(s + i, w, h, 0, 5, d - 1, 8) for i, ((w, h), d) in enumerate(zip(sizes, disps))]
            for s in (30, 40)]
    want = [[_want(cs.ALGO_CENSUS, L, R, d, (9, 9), (9, 7)) for (L, R), d in zip(pairs, disps)] for pairs in sets]
    b = cs.BatchMatcher(ctx, sizes, disps, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7))
    b.upload(sets[0])
    for f in range(NFRAMES):
        k = f % 2
        b.refill(sets[k])
        b.process()
        for i in range(len(sizes)):
            got = (b.outs[i].download(), b.lrs[i].download(), b.rls[i].download())
            for a, c in zip(got, want[k][i]):
                assert np.array_equal(a, c), f"frame {f} (set {k}, pair {i})"
