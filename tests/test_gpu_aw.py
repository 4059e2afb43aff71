"""GPU adaptive-window matchers (AW1 / AW2) against the CPU oracle: bit-exact maps
and bit-exact integer window values."""
import numpy as np
import pytest

import clutec_stereo as cs
import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo,mode", [(cs.ALGO_AW2, 2), (cs.ALGO_AW1, 1)])
@pytest.mark.parametrize("W,H,D,sub,kind", [
    (200, 100, 32, 5, 1),
    (257, 131, 64, 3, 1),
    (190, 97, 40, 5, 0),
])
def test_aw_bit_exact(ctx, algo, mode, W, H, D, sub, kind):
    print(f"aw algo={algo} {W}x{H} D={D} sub={sub}", flush=True)
    L, R = cs.synth_pair(W + H + D, W, H, kind, 6, D - 1, min(8, D - 1))
    m = cs.Matcher(ctx, W, H, algorithm=algo, window=(sub, sub), num_disp=D, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ids = m.run(L, R, dispid=True)
    ref_out, ref_l, ref_r, val = orc.aw(mode, L, R, D, sub=sub, tol=1, nthreads=8)
    assert np.array_equal(lr, ref_l)
    assert np.array_equal(rl, ref_r)
    assert np.array_equal(out, ref_out)
    v = ref_l >= 0
    # channel 1 of the reference pixel: AW1 similarity / AW2 SAD of the winning window
    assert np.array_equal(ids[:, :, 1][v].astype(np.int64), val[v].astype(np.int64))


def test_aw_config2_full(ctx):
    """1280x720 slanted planes, D=128, 15x15 window of 5x5 sub-patches."""
    L, R = cs.synth_pair(2, 1280, 720, 1, 6, 127, 8)
    m = cs.Matcher(ctx, 1280, 720, algorithm=cs.ALGO_AW2, window=(5, 5), num_disp=128, tol=1)
    out, lr, rl = m.run(L, R, maps=True)
    ref_out, ref_l, ref_r, _ = orc.aw(2, L, R, 128, sub=5, tol=1, nthreads=16)
    assert np.array_equal(lr, ref_l) and np.array_equal(rl, ref_r) and np.array_equal(out, ref_out)
