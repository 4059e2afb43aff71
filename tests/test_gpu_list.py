"""Work-list matcher: the batched mode (many pairs of uneven size in one persistent,
graph-captured launch) and the band search of the learned proposal, each bit-exact
against the one-pair matcher and the oracle."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu

SIZES = [(320, 240, 64), (256, 160, 128), (400, 208, 192), (192, 144, 64), (512, 96, 256), (272, 272, 128)]


def _pairs(seed0):
    return [cs.synth_pair(seed0 + i, w, h, i % 3, 6, d - 1, min(8, d - 1)) for i, (w, h, d) in enumerate(SIZES)]


def _singles(ctx, pairs, algo, window, census):
    out = []
    for (w, h, d), (L, R) in zip(SIZES, pairs):
        m = cs.Matcher(ctx, w, h, algorithm=algo, window=window, census=census, num_disp=d, tol=1)
        out.append(m.run(L, R, maps=True))
    return out


@pytest.mark.parametrize("algo,window,census", [(cs.ALGO_SAD, (7, 7), (1, 1)), (cs.ALGO_CENSUS, (9, 9), (9, 7))])
def test_batch_matches_single_pairs(ctx, algo, window, census):
    print(f"batch algo={algo}", flush=True)
    pairs = _pairs(500)
    b = cs.BatchMatcher(ctx, [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES], algorithm=algo, window=window,
                        census=census)
    got = b.run(pairs)
    want = _singles(ctx, pairs, algo, window, census)
    for i, (g, w) in enumerate(zip(got, want)):
        for a, c in zip(g, w):
            assert np.array_equal(a, c), f"pair {i}"
    # and one pair against the oracle directly
    (w0, h0, d0), (L, R) = SIZES[1], pairs[1]
    if algo == cs.ALGO_CENSUS:
        ref = orc.bm(orc.COST_CENSUS, orc.census(L, 9, 7), orc.census(R, 9, 7), d0, window=window, census_win=(9, 7),
                     tol=1, nthreads=8)
    else:
        ref = orc.bm(orc.COST_SAD, L, R, d0, window=window, tol=1, nthreads=8)
    for a, c in zip(got[1], ref):
        assert np.array_equal(a, c)


def test_batch_graph_replays(ctx):
    """Replays of the captured graph, with eager work and new content in between,
    each equal to fresh one-pair results."""
    algo, window, census = cs.ALGO_CENSUS, (9, 9), (9, 7)
    b = cs.BatchMatcher(ctx, [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES], algorithm=algo, window=window,
                        census=census, tuning={"chunk": 64, "grid": 512, "order": 0, "use_graph": 1})
    first = _pairs(600)
    b.upload(first)
    for rep, seed in enumerate((600, 700, 600)):
        pairs = _pairs(seed)
        b.refill(pairs)
        b.process()
        got = b.results()
        want = _singles(ctx, pairs, algo, window, census)  # eager work between the replays
        for g, w in zip(got, want):
            for a, c in zip(g, w):
                assert np.array_equal(a, c), f"replay {rep}"


@pytest.mark.parametrize("tuning", [{"chunk": 32, "grid": 256, "order": 1, "use_graph": 0},
                                    {"chunk": 256, "grid": 768, "order": 0, "use_graph": 1},
                                    {"chunk": 96, "grid": 300, "order": 0, "use_graph": 0}])
def test_batch_tunings_agree(ctx, tuning):
    algo, window, census = cs.ALGO_SAD, (7, 7), (1, 1)
    pairs = _pairs(800)
    b = cs.BatchMatcher(ctx, [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES], algorithm=algo, window=window,
                        census=census, tuning=tuning)
    got = b.run(pairs)
    want = _singles(ctx, pairs, algo, window, census)
    for g, w in zip(got, want):
        for a, c in zip(g, w):
            assert np.array_equal(a, c)


def test_batch_autotune(ctx):
    pairs = _pairs(900)
    b = cs.BatchMatcher(ctx, [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES], algorithm=cs.ALGO_CENSUS,
                        window=(9, 9), census=(9, 7))
    b.upload(pairs)
    t = b.autotune()
    print(f"autotuned {t}", flush=True)
    assert t["chunk"] in (32, 64, 128, 256) and t["order"] in (0, 1) and t == b.tuning()
    b.process()
    got = b.results()
    want = _singles(ctx, pairs, cs.ALGO_CENSUS, (9, 9), (9, 7))
    for g, w in zip(got, want):
        for a, c in zip(g, w):
            assert np.array_equal(a, c)


def test_batch_rejects_bad_config(ctx):
    with pytest.raises(cs.StereoError) as e:
        cs.BatchMatcher(ctx, [(320, 240)], [64], algorithm=cs.ALGO_NCC, window=(9, 9))
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    with pytest.raises(cs.StereoError) as e:
        cs.BatchMatcher(ctx, [(320, 240)], [60], algorithm=cs.ALGO_SAD, window=(7, 7))
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    with pytest.raises(cs.StereoError) as e:
        cs.BatchMatcher(ctx, [(320, 240)], [64], algorithm=cs.ALGO_SAD, window=(7, 7), tuning={"chunk": 12})
    assert e.value.status == cs.ERROR_INVALID_CONFIG


def test_learned_full_fallback_equals_full_range(ctx):
    """With every tile below the confidence threshold the learned matcher is the
    full-range census matcher."""
    W, H, D = 384, 224, 128
    L, R = cs.synth_pair(77, W, H, 2, 8, D - 1, 8)
    m = cs.LearnedMatcher(ctx, W, H, cs.cnn_init_params(0), num_disp=D, logit_thresh=1e30)
    got = m.run(L, R, maps=True)
    st = m.stats()
    assert st["fallback_tiles"] == st["tiles"]
    full = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1)
    want = full.run(L, R, maps=True)
    for a, c in zip(got, want):
        assert np.array_equal(a, c)


def test_learned_band_matches_oracle(ctx):
    """Band search (no fallback): every tile searches only its proposed disparity
    bins; the maps equal the oracle restricted to the same per-tile bins."""
    W, H, D = 384, 224, 128
    L, R = cs.synth_pair(78, W, H, 2, 8, D - 1, 8)
    m = cs.LearnedMatcher(ctx, W, H, cs.cnn_init_params(0), num_disp=D, logit_thresh=-1e30, chunk=32)
    got = m.run(L, R, maps=True)
    masks = m.tile_masks()
    st = m.stats()
    assert st["fallback_tiles"] == 0
    bins = np.repeat(np.repeat(masks, 56, axis=0), 64, axis=1)[:H, :W]
    assert (masks != 0).all() and (masks < (1 << (D // 8))).all()
    nbits = np.array([bin(int(v)).count("1") for v in bins.ravel()]).reshape(H, W)
    assert st["evals"] == int((8 * nbits).astype(np.int64).sum())
    ref = orc.bm_bins(orc.COST_CENSUS, orc.census(L, 9, 7), orc.census(R, 9, 7), D, bins, window=(9, 9),
                      census_win=(9, 7), tol=1, nthreads=8)
    for a, c in zip(got, ref):
        assert np.array_equal(a, c)


def test_learned_graph_replays(ctx):
    """Untimed calls replay a captured graph; replays with new content in the same
    input buffers, and a timed direct run, all give the same maps as fresh runs."""
    W, H, D = 384, 224, 128
    params = cs.cnn_init_params(3)
    m = cs.LearnedMatcher(ctx, W, H, params, num_disp=D, logit_thresh=-2.0)
    dl = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U8)
    dr = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U8)
    out = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
    for rep, seed in enumerate((80, 81, 80)):
        L, R = cs.synth_pair(seed, W, H, 2, 8, D - 1, 8)
        dl.upload(L)
        dr.upload(R)
        m.process(out, None, None, dl, dr)
        got = out.download()
        fresh = cs.LearnedMatcher(ctx, W, H, params, num_disp=D, logit_thresh=-2.0)
        ctx.set_timing(True)
        want = fresh.run(L, R)
        ctx.set_timing(False)
        assert np.array_equal(got, want), f"replay {rep}"
