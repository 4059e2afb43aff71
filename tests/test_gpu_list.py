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


def test_batch_autotune(ctx, tmp_path):
    """The autotuner's per-class table is persisted, applied to a second driver of the
    same pairs, and both give the one-pair results."""
    pairs = _pairs(900)
    sizes, ds = [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES]
    b = cs.BatchMatcher(ctx, sizes, ds, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7))
    b.upload(pairs)
    path = tmp_path / "tuning.json"
    t = b.autotune(table_path=str(path))
    table = b.tuning_table()
    print(f"autotuned {t}, table {table}", flush=True)
    assert t["chunk"] in (32, 64, 128, 256) and t["order"] in (0, 1) and t == b.tuning()
    assert table["matcher"] == "census" and table["chunk"] == t["chunk"] and table["grid"] == t["grid"]
    # every pair here has under 64 tiles: one class per disparity count
    assert sorted(c["disparities"] for c in table["classes"]) == sorted(set(ds))
    assert all(c["tiles"] == "small" and c["tile"] == "64x56" and c["chunk"] in (32, 64, 128, 256)
               for c in table["classes"])
    b.process()
    got = b.results()
    want = _singles(ctx, pairs, cs.ALGO_CENSUS, (9, 9), (9, 7))
    for g, w in zip(got, want):
        for a, c in zip(g, w):
            assert np.array_equal(a, c)
    b2 = cs.BatchMatcher(ctx, sizes, ds, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7))
    b2.apply_tuning_table(str(path))
    assert b2.tuning_table() == table
    for g, w in zip(b2.run(pairs), want):
        for a, c in zip(g, w):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("mix", [{64: 64, 128: 128, 192: 96, 256: 128}, {64: 32, 128: 128, 192: 192, 256: 256}])
def test_batch_per_class_items(ctx, mix):
    """Pairs of one batch cut into work items of different lengths per disparity count:
    bit-exact against the one-pair matcher."""
    algo, window, census = cs.ALGO_SAD, (7, 7), (1, 1)
    pairs = _pairs(1000)
    b = cs.BatchMatcher(ctx, [(w, h) for w, h, _ in SIZES], [d for _, _, d in SIZES], algorithm=algo, window=window,
                        census=census, tuning={"chunk": 64, "grid": 512, "order": 0, "use_graph": 1})
    table = b.tuning_table()
    table["classes"] = [{"disparities": d, "tiles": "small", "chunk": c} for d, c in mix.items()]
    b.apply_tuning_table(table)
    got_table = b.tuning_table()
    assert {c["disparities"]: c["chunk"] for c in got_table["classes"]} == mix
    got = b.run(pairs)
    want = _singles(ctx, pairs, algo, window, census)
    for i, (g, w) in enumerate(zip(got, want)):
        for a, c in zip(g, w):
            assert np.array_equal(a, c), f"pair {i}"


def test_batch_refuses_bad_tables(ctx):
    b = cs.BatchMatcher(ctx, [(320, 240), (256, 160)], [64, 128], algorithm=cs.ALGO_SAD, window=(7, 7))
    keep = b.tuning_table()
    bad = [dict(keep, chunk=12), dict(keep, classes=[{"disparities": 64, "tiles": "huge", "chunk": 64}]),
           dict(keep, classes=[{"disparities": 64, "chunk": 64}]),
           dict(keep, classes=[{"disparities": 64, "tiles": "small", "chunk": 300}]), dict(keep, grid=0)]
    for t in bad:
        with pytest.raises(cs.StereoError) as e:
            b.apply_tuning_table(t)
        assert e.value.status == cs.ERROR_INVALID_CONFIG
        assert b.tuning_table() == keep


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


@pytest.mark.parametrize("centre_residual", [False, True])
def test_learned_band_matches_oracle(ctx, centre_residual):
    """Band search (no fallback): every tile searches only its proposed disparity
    bins; the maps equal the oracle restricted to the same per-tile bins."""
    W, H, D = 384, 224, 128
    L, R = cs.synth_pair(78, W, H, 2, 8, D - 1, 8)
    m = cs.LearnedMatcher(ctx, W, H, cs.cnn_init_params(0), num_disp=D, logit_thresh=-1e30, chunk=32,
                          centre_residual=centre_residual)
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


def _coarse_fill(dc):
    """Holes of a quarter-resolution map filled with the smaller of the nearest valid
    values within 64 pixels to the left and right along the row (0 when none)."""
    out = dc.astype(np.int64).copy()
    H4, W4 = dc.shape
    for y in range(H4):
        for x in range(W4):
            if dc[y, x] >= 0:
                continue
            left = [dc[y, i] for i in range(x - 1, max(-1, x - 65), -1) if dc[y, i] >= 0][:1]
            right = [dc[y, i] for i in range(x + 1, min(W4, x + 65)) if dc[y, i] >= 0][:1]
            cand = left + right
            out[y, x] = min(cand) if cand else 0
    return out


def test_learned_masks_from_coarse_map(ctx):
    """The learned matcher's tile bins equal the bins built on the CPU: 4 x 4 box
    means (sum + 8) >> 4 of both images, the oracle's census matcher at quarter
    resolution (D/4), the row-wise hole fill, and every quarter pixel of the tile and
    its one-pixel margin proposing [4 c - 8, 4 c + 8] (band 8)."""
    W, H, D = 384, 224, 128
    L, R = cs.synth_pair(79, W, H, 1, 8, D - 1, 8)
    m = cs.LearnedMatcher(ctx, W, H, cs.cnn_init_params(0), num_disp=D, logit_thresh=-1e30, band=8)
    m.run(L, R)
    got = m.tile_masks()

    def ds4(img):
        return ((img.reshape(H // 4, 4, W // 4, 4).astype(np.int32).sum(axis=(1, 3)) + 8) >> 4).astype(np.uint8)

    L4, R4 = ds4(L), ds4(R)
    dc = orc.bm(orc.COST_CENSUS, orc.census(L4, 9, 7), orc.census(R4, 9, 7), D // 4, window=(9, 9),
                census_win=(9, 7), tol=1)[0]
    c = 4 * _coarse_fill(dc)
    W4, H4 = W // 4, H // 4
    ntx, nty = (W + 63) // 64, (H + 55) // 56
    want = np.zeros((nty, ntx), np.uint64)
    for ty in range(nty):
        for tx in range(ntx):
            blk = c[max(0, 14 * ty - 1):min(H4, 14 * ty + 15), max(0, 16 * tx - 1):min(W4, 16 * tx + 17)]
            mask = 0
            for v in np.unique(blk):
                lo, hi = max(0, int(v) - 8), min(D - 1, int(v) + 8)
                for b in range(lo >> 3, (hi >> 3) + 1):
                    mask |= 1 << b
            want[ty, tx] = mask
    assert np.array_equal(got.astype(np.uint64), want)


def test_learned_graph_replays(ctx):
    """With graph_replay, untimed calls replay a captured graph; replays with new content
    in the same input buffers, a timed run and the default direct launch all give the
    same maps as fresh runs."""
    W, H, D = 384, 224, 128
    params = cs.cnn_init_params(3)
    m = cs.LearnedMatcher(ctx, W, H, params, num_disp=D, logit_thresh=-2.0, graph_replay=True)
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
        direct = cs.LearnedMatcher(ctx, W, H, params, num_disp=D, logit_thresh=-2.0)
        assert np.array_equal(direct.run(L, R), want), f"direct launch {rep}"
