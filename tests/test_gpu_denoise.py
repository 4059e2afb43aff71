"""Disparity de-noise (Statistics::DispDenoise) on the GPU against the CPU oracle, bit-exact."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu


def _noisy_map(seed, W, H, hole=0.2, spike=0.05):
    rng = np.random.default_rng(seed)
    base = (60 + 40 * np.sin(np.arange(W)[None, :] / 37.0) + np.arange(H)[:, None] / 9.0).astype(np.int64)
    d = base + rng.integers(-1, 2, (H, W))
    d[rng.random((H, W)) < hole] = -1
    d[rng.random((H, W)) < spike] = rng.integers(0, 250)
    return d.astype(np.int16)


def _as_ref(d, seed):
    rng = np.random.default_rng(seed)
    arr = np.zeros(d.shape + (4,), np.uint16)
    arr[..., 0] = np.where(d >= 0, d + 1, np.where(rng.random(d.shape) < 0.5, 0, 0xFFFA))
    arr[..., 1] = 7  # other channels carry nothing the filter reads
    return arr


@pytest.mark.parametrize("P", [3, 5, 7, 9, 11, 16])
@pytest.mark.parametrize("fmt", ["i16", "ref"])
def test_denoise_matches_oracle(ctx, P, fmt):
    W, H = 331, 197
    print(f"denoise P {P} {fmt} {W}x{H}", flush=True)
    d = _noisy_map(P, W, H)
    arr = _as_ref(d, P) if fmt == "ref" else d
    for ratio, cutoff, fill in [(0.5, 1.0, True), (0.3, 2.5, False), (0.8, 0.75, True)]:
        dn = cs.Denoise(ctx, W, H, patch=P, support_ratio=ratio, mean_cutoff=cutoff, fill=fill,
                        ref_format=(fmt == "ref"))
        got = dn.run(arr)
        want = orc.disp_denoise(arr, P=P, support_ratio=ratio, mean_cutoff=cutoff, fill=fill)
        assert got.dtype == want.dtype
        assert np.array_equal(got, want), (ratio, cutoff, fill, int((got != want).sum()))


@pytest.mark.parametrize("size", [(16, 16), (17, 40), (64, 16), (65, 17), (128, 200)])
def test_denoise_edge_sizes(ctx, size):
    W, H = size
    P = 16 if min(W, H) >= 16 else 9
    print(f"denoise edge {W}x{H} P {P}", flush=True)
    d = _noisy_map(W * 1000 + H, W, H)
    dn = cs.Denoise(ctx, W, H, patch=P)
    assert np.array_equal(dn.run(d), orc.disp_denoise(d, P=P))


def test_denoise_matcher_output_1080p(ctx):
    W, H, D = 1920, 1080, 256
    print("denoise of a 1080p census matcher map", flush=True)
    L, R = cs.synth_pair(5, W, H, 0, 8, D - 1, 8)
    m = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, num_disp=D, tol=1)
    disp = m.run(L, R)
    for fmt in ["i16", "ref"]:
        arr = _as_ref(disp, 3) if fmt == "ref" else disp
        dn = cs.Denoise(ctx, W, H, patch=5, ref_format=(fmt == "ref"))
        got = dn.run(arr)
        want = orc.disp_denoise(arr, P=5)
        assert np.array_equal(got, want), fmt
    # the filter removes isolated values: no output exceeds the range of its input
    assert got.max() <= arr[..., 0].max()


def test_denoise_configuration_errors(ctx):
    with pytest.raises(cs.StereoError) as e:
        cs.Denoise(ctx, 64, 64, patch=4)
    assert e.value.status == cs.ERROR_INVALID_CONFIG
    with pytest.raises(cs.StereoError) as e:
        cs.Denoise(ctx, 10, 64, patch=11)
    assert e.value.status == cs.ERROR_INVALID_FORMAT
    dn = cs.Denoise(ctx, 64, 48, patch=5)
    src = cs.DeviceImage.from_array(ctx, np.zeros((48, 64), np.int16), cs.PIXEL_LUM_I16)
    bad = cs.DeviceImage(ctx, 64, 48, cs.PIXEL_LUM_U16)
    with pytest.raises(cs.StereoError):
        dn.process(bad, src)
