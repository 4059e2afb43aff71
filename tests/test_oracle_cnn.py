"""The CPU oracle of the proposal network, pinned without the GPU: the convolution
against an independent numpy formulation, its gradients against central finite
differences in FP64, and bf16 rounding against its definition."""
import numpy as np
import pytest

import cnn_ref
import orc


def _np_conv(x, w, b, dil):
    N, H, W, _ = x.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (dil, dil), (dil, dil), (0, 0)))
    y = np.zeros((N, H, W, w.shape[0])) + (0 if b is None else b)
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("nhwc,oc->nhwo", xp[:, ky * dil:ky * dil + H, kx * dil:kx * dil + W],
                           w[:, ky, kx, :].astype(np.float64))
    return y


@pytest.mark.parametrize("dil", [1, 2, 4])
def test_conv_matches_numpy(dil):
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((2, 9, 13, 5)).astype(np.float32)
    w = rng.standard_normal((3, 3, 3, 5)).astype(np.float32)
    b = rng.standard_normal(3).astype(np.float32)
    assert np.abs(orc.conv3x3(x, w, b, dil) - _np_conv(x, w, b, dil)).max() < 1e-5


def test_conv_impulse_picks_flipped_tap():
    """A unit impulse at (y, x) lands output tap (ky, kx) at (y - (ky-1) r, x - (kx-1) r)."""
    x = np.zeros((1, 11, 11, 1), np.float32)
    x[0, 5, 5, 0] = 1.0
    w = np.arange(9, dtype=np.float32).reshape(1, 3, 3, 1) + 1
    y = orc.conv3x3(x, w, None, 2)[0, :, :, 0]
    for ky in range(3):
        for kx in range(3):
            assert y[5 - (ky - 1) * 2, 5 - (kx - 1) * 2] == w[0, ky, kx, 0]
    assert np.count_nonzero(y) == 9


@pytest.mark.parametrize("dil", [1, 3])
def test_conv_gradients_finite_differences(dil):
    rng = np.random.default_rng(10 + dil)
    x = rng.standard_normal((1, 6, 7, 3)).astype(np.float32)
    w = rng.standard_normal((2, 3, 3, 3)).astype(np.float32)
    b = rng.standard_normal(2).astype(np.float32)
    r = rng.standard_normal((1, 6, 7, 2))  # loss = sum(r * y)
    dx, dw, db = orc.conv3x3_bwd(x, w, r.astype(np.float32), dil)
    f = lambda xx, ww, bb: float((r * _np_conv(xx, ww, bb, dil)).sum())
    h = 1e-3
    for idx in [(0, 0, 0, 0), (0, 3, 4, 1), (0, 5, 6, 2)]:
        xp, xm = x.astype(np.float64).copy(), x.astype(np.float64).copy()
        xp[idx] += h
        xm[idx] -= h
        assert abs((f(xp, w, b) - f(xm, w, b)) / (2 * h) - dx[idx]) < 1e-3
    for idx in [(0, 0, 0, 0), (1, 2, 1, 2), (0, 1, 2, 1)]:
        wp, wm = w.astype(np.float64).copy(), w.astype(np.float64).copy()
        wp[idx] += h
        wm[idx] -= h
        assert abs((f(x, wp, b) - f(x, wm, b)) / (2 * h) - dw[idx]) < 1e-3
    assert np.allclose(db, r.sum(axis=(0, 1, 2)))


def test_bf16_round_definition():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e38], np.float32)
    r = orc.bf16_round(v)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5 and r[5] == 0.0
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_network_gradient_finite_differences():
    """The network backward (no rounding) against central differences of the loss."""
    rng = np.random.default_rng(3)
    p = (rng.standard_normal(cnn_ref.NPARAMS) * 0.2).astype(np.float32)
    x0 = rng.random((1, 8, 10, 2)).astype(np.float32)
    label = (rng.random((1, 8, 10)) * 4).astype(np.float32)
    label[0, 0, :3] = np.nan
    dfill = np.zeros((1, 8, 10), np.float32)
    out, acts = cnn_ref.forward(p, x0, rounding=False)
    g, _ = cnn_ref.loss_grad(out, label, dfill)
    g[..., 1] = 0.0  # Huber part only: the confidence target is piecewise constant
    grads = cnn_ref.backward(p, acts, g, rounding=False)
    h = 2e-3
    p64 = p.astype(np.float64)
    for k in [0, 17, 300, 1000, 3000, 6000, cnn_ref.NPARAMS - 3, cnn_ref.NPARAMS - 1]:
        pp, pm = p64.copy(), p64.copy()
        pp[k] += h
        pm[k] -= h
        num = (cnn_ref.loss_value(pp.astype(np.float32), x0, label, dfill) -
               cnn_ref.loss_value(pm.astype(np.float32), x0, label, dfill)) / (2 * h)
        assert abs(num - grads[k]) <= 2e-3 * max(1.0, abs(grads[k])) + 2e-4, (k, num, grads[k])
