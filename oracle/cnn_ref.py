"""FP64 restatement of the learned proposal network -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ (and the smoke check) as the checker of the GPU network; the
engine never uses it.  The layer stack, the bf16 rounding points and the loss
follow the GPU implementation's definition (module.clutec.cuda_amd/csrc/cnn.hip,
learned.cpp):

  a0 = bf16(input)                              (N, H, W, 2)
  z_l = conv3x3(a_l, bf16(w_l), dil_l) + b_l     FP64 accumulation
  a_{l+1} = bf16(relu(z_l))                     l = 0..3
  out = z_4                                     (centre residual, confidence logit)
  loss = (1/P) [sum over pixels with a disparity label of Huber(e)
                + lambda * sum over pixels with a confidence target of BCE(out1, c)],
         e = out0 - (label - dfill), c = the confidence label where given, else |e| <= tau
  backward: g4 = bf16(dloss/dout); dW_l, db_l from (a_l, g_l);
            g_{l-1} = bf16(conv3x3_bwd_data(g_l, bf16(w_l)) * (a_l > 0))
The convolution and its gradients are the C oracle's (cnn_oracle.c).
"""
import numpy as np

import orc

LAYERS = [(2, 16, 1), (16, 16, 2), (16, 16, 4), (16, 16, 2), (16, 2, 1)]
NPARAMS = sum(co * 9 * ci + co for ci, co, _ in LAYERS)
LAMBDA = 0.5
TAU = 2.0


def unpack(params):
    """Public flat layout -> [(w [cout,3,3,cin], b [cout])]."""
    params = np.asarray(params, np.float32)
    assert params.size == NPARAMS
    out, k = [], 0
    for ci, co, _ in LAYERS:
        w = params[k:k + co * 9 * ci].reshape(co, 3, 3, ci)
        k += co * 9 * ci
        b = params[k:k + co]
        k += co
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([np.asarray(w, np.float64).ravel(), np.asarray(b, np.float64).ravel()])
                           for w, b in layers])


def forward(params, x0, rounding=True):
    """x0 (N,H,W,2) float -> (out (N,H,W,2) float32, activations a_0..a_4)."""
    rnd = orc.bf16_round if rounding else (lambda a: np.asarray(a, np.float32))
    a = rnd(x0)
    acts = [a]
    out = None
    for l, ((ci, co, dil), (w, b)) in enumerate(zip(LAYERS, unpack(params))):
        z = orc.conv3x3(a, rnd(w), b, dil)
        if l + 1 < len(LAYERS):
            a = rnd(np.maximum(z, 0.0))
            acts.append(a)
        else:
            out = z
    return out, acts


def loss_grad(out, label, dfill, clabel=None):
    """Loss sums and dloss/dout (FP32) with the GPU's normalisation 1/P.  clabel:
    optional confidence targets (0 or 1, NaN = none) that replace |e| <= tau."""
    P = out.shape[0] * out.shape[1] * out.shape[2]
    o0 = out[..., 0].astype(np.float64)
    o1 = out[..., 1].astype(np.float64)
    t = label.astype(np.float64) - dfill.astype(np.float64)
    lab = np.isfinite(t)
    e = np.where(lab, o0 - np.where(lab, t, 0.0), 0.0)
    ae = np.abs(e)
    c = np.where(lab, (ae <= TAU).astype(np.float64), -1.0)
    if clabel is not None:
        cl = np.asarray(clabel, np.float64)
        c = np.where(np.isfinite(cl), cl, c)
    has_c = c >= 0
    sg = 1.0 / (1.0 + np.exp(-o1))
    g = np.zeros(out.shape, np.float64)
    g[..., 0] = np.where(lab, np.clip(e, -1, 1) / P, 0.0)
    g[..., 1] = np.where(has_c, LAMBDA * (sg - c) / P, 0.0)
    hub = np.where(ae <= 1, 0.5 * e * e, ae - 0.5)
    bce = np.maximum(o1, 0) - o1 * c + np.log1p(np.exp(-np.abs(o1)))
    n = (lab | has_c).sum()
    return g.astype(np.float32), {"huber": float(hub[lab].sum() / max(n, 1)), "bce": float(bce[has_c].sum() / max(n, 1)),
                                  "count": int(n)}


def backward(params, acts, gout, rounding=True):
    """Gradients of every parameter (public layout, FP64) for dloss/dout = gout."""
    rnd = orc.bf16_round if rounding else (lambda a: np.asarray(a, np.float32))
    layers = unpack(params)
    g = rnd(gout)
    grads = [None] * len(LAYERS)
    for l in range(len(LAYERS) - 1, -1, -1):
        ci, co, dil = LAYERS[l]
        w, b = layers[l]
        dx, dw, db = orc.conv3x3_bwd(acts[l], rnd(w), g, dil, want_dx=l > 0)
        grads[l] = (dw, db)
        if l > 0:
            g = rnd(dx * (acts[l] > 0))
    return pack(grads)


def loss_value(params, x0, label, dfill):
    """Scalar loss with the GPU's normalisation (sum over labelled / P), no rounding."""
    out, _ = forward(params, x0, rounding=False)
    P = out.shape[0] * out.shape[1] * out.shape[2]
    o0 = out[..., 0].astype(np.float64)
    o1 = out[..., 1].astype(np.float64)
    t = label.astype(np.float64) - dfill.astype(np.float64)
    lab = np.isfinite(t)
    e = o0 - np.where(lab, t, 0.0)
    ae = np.abs(e)
    hub = np.where(ae <= 1, 0.5 * e * e, ae - 0.5)
    return float(hub[lab].sum() / P)
