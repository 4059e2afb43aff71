"""GPU kernels of the proposal network against the FP64 oracle.

Convolutions run on bf16 MFMA with FP32 accumulation, so on bf16-exact inputs the
FP32 outputs agree with the FP64 oracle to FP32 accumulation error (checked
relative to the sum of absolute products); bf16 outputs agree to one bf16 rounding.
Network gradients are checked layer by layer against the oracle run with the same
bf16 rounding points."""
import ctypes as C

import numpy as np
import pytest

import clutec_stereo as cs
import cnn_ref
import orc

pytestmark = pytest.mark.gpu

FP32_RTOL = 2e-6  # of sum |w x|: FP32 accumulation over <= 288 products


def _bf16_bits(a):
    r = orc.bf16_round(a)
    return r, (r.view(np.uint32) >> 16).astype(np.uint16)


def _conv(ctx, x, w, b, dil, relu):
    N, H, W, ci = x.shape
    co = w.shape[0]
    xq, xb = _bf16_bits(x)
    wq, wb = _bf16_bits(w)
    tx = cs.DeviceBuffer.from_array(ctx, xb)
    tw = cs.DeviceBuffer.from_array(ctx, wb.reshape(co, 9, ci))
    tb = cs.DeviceBuffer.from_array(ctx, b.astype(np.float32)) if b is not None else None
    ty = cs.DeviceBuffer(ctx, N * H * W * co * (2 if relu else 4))
    st = cs.lib.stereo_conv3x3_bf16(ctx.handle, tx.ptr, N, H, W, ci, tw.ptr, tb.ptr if tb is not None else None, co,
                                    dil, 1 if relu else 0, ty.ptr, None)
    cs._check(st, ctx)
    cs._check(cs.lib.stereo_synchronize(ctx.handle, None), ctx)
    if relu:
        y = ty.download(np.uint16, (N, H, W, co))
        y = (y.astype(np.uint32) << 16).view(np.float32)
    else:
        y = ty.download(np.float32, (N, H, W, co))
    return xq, wq, y


@pytest.mark.parametrize("ci,co,dil", [(16, 32, 1), (32, 32, 2), (32, 32, 4), (32, 16, 1)])
def test_conv_forward_fp32(ctx, ci, co, dil):
    print(f"conv fwd {ci}->{co} dil {dil}", flush=True)
    rng = np.random.default_rng(ci * 7 + co + dil)
    x = rng.standard_normal((2, 13, 37, ci)).astype(np.float32)
    w = (rng.standard_normal((co, 3, 3, ci)) * 0.1).astype(np.float32)
    b = rng.standard_normal(co).astype(np.float32)
    xq, wq, y = _conv(ctx, x, w, b, dil, relu=False)
    ref = orc.conv3x3(xq, wq, b, dil).astype(np.float64)
    mag = orc.conv3x3(np.abs(xq), np.abs(wq), np.abs(b), dil).astype(np.float64)
    err = np.abs(y - ref) / np.maximum(mag, 1e-30)
    print(f"  max error / sum|terms| = {err.max():.2e}", flush=True)
    assert err.max() <= FP32_RTOL


@pytest.mark.parametrize("ci,co,dil", [(16, 32, 1), (32, 32, 2)])
def test_conv_forward_relu_bf16(ctx, ci, co, dil):
    rng = np.random.default_rng(5 + dil)
    x = rng.standard_normal((1, 17, 45, ci)).astype(np.float32)
    w = (rng.standard_normal((co, 3, 3, ci)) * 0.1).astype(np.float32)
    b = rng.standard_normal(co).astype(np.float32)
    xq, wq, y = _conv(ctx, x, w, b, dil, relu=True)
    ref = np.maximum(orc.conv3x3(xq, wq, b, dil), 0.0)
    assert (y >= 0).all()
    # one bf16 rounding of a value that carries FP32 accumulation error
    assert np.all(np.abs(y - ref) <= np.abs(ref) * 2.0 ** -8 + 1e-6)
    assert np.mean(y == orc.bf16_round(ref)) > 0.999


@pytest.mark.parametrize("ci,co,dil", [(16, 32, 1), (32, 32, 4), (32, 16, 2)])
def test_conv_backward(ctx, ci, co, dil):
    print(f"conv bwd {ci}->{co} dil {dil}", flush=True)
    rng = np.random.default_rng(ci + 3 * co + dil)
    N, H, W = 2, 11, 29
    xq, xb = _bf16_bits(rng.standard_normal((N, H, W, ci)).astype(np.float32))
    wq, wb = _bf16_bits((rng.standard_normal((co, 3, 3, ci)) * 0.1).astype(np.float32))
    gq, gb = _bf16_bits(rng.standard_normal((N, H, W, co)).astype(np.float32))
    wt = np.ascontiguousarray(wb.reshape(co, 9, ci).transpose(2, 1, 0))  # [ci][9][co]
    tx, twt, tg = (cs.DeviceBuffer.from_array(ctx, a) for a in (xb, wt, gb))
    tdx = cs.DeviceBuffer(ctx, N * H * W * ci * 4)
    tdw = cs.DeviceBuffer(ctx, (co * 9 * ci + co) * 4)
    cs._check(cs.lib.stereo_conv3x3_bwd_bf16(ctx.handle, tx.ptr, N, H, W, ci, twt.ptr, tg.ptr, co, dil, tdx.ptr, tdw.ptr,
                                             None), ctx)
    cs._check(cs.lib.stereo_synchronize(ctx.handle, None), ctx)
    dx = tdx.download(np.float32, (N, H, W, ci))
    dwdb = tdw.download(np.float32, (co * 9 * ci + co,))
    rdx, rdw, rdb = orc.conv3x3_bwd(xq, wq, gq, dil)
    mdx, mdw, mdb = orc.conv3x3_bwd(np.abs(xq), np.abs(wq), np.abs(gq), dil)
    e_dx = (np.abs(dx - rdx) / np.maximum(mdx, 1e-30)).max()
    dw = dwdb[:co * 9 * ci].reshape(co, 3, 3, ci)
    e_dw = (np.abs(dw - rdw) / np.maximum(mdw, 1e-30)).max()
    e_db = (np.abs(dwdb[co * 9 * ci:] - rdb) / np.maximum(mdb, 1e-30)).max()
    print(f"  dx {e_dx:.2e} dw {e_dw:.2e} db {e_db:.2e} (of sum|terms|)", flush=True)
    assert e_dx <= FP32_RTOL and e_dw <= 1e-5 and e_db <= 1e-5


def _fill(dc):
    """The coarse-disparity hole filling of the network input (nearest valid to the
    left and right within 64 pixels, the smaller; 0 when none)."""
    H, W = dc.shape
    out = dc.astype(np.int32).copy()
    for y in range(H):
        for x in range(W):
            if dc[y, x] >= 0:
                continue
            a = b = -1
            for k in range(1, 65):
                if a < 0 and x - k >= 0 and dc[y, x - k] >= 0:
                    a = dc[y, x - k]
                if b < 0 and x + k < W and dc[y, x + k] >= 0:
                    b = dc[y, x + k]
                if a >= 0 and b >= 0:
                    break
            out[y, x] = 0 if (a < 0 and b < 0) else (b if a < 0 else (a if b < 0 else min(a, b)))
    return out


def _labels(disp, W4, H4):
    lab = np.full((H4, W4), np.nan, np.float32)
    for Y in range(H4):
        for X in range(W4):
            blk = disp[4 * Y:4 * Y + 4, 4 * X:4 * X + 4].astype(np.int64)
            v = blk[blk >= 0]
            if v.size >= 8 and v.max() - v.min() <= 4:
                lab[Y, X] = np.float32(v.sum()) / np.float32(4 * v.size)
    return lab


def _frame(ctx, seed, W, H, D):
    """One training frame: quarter-resolution inputs, coarse map and labels, built
    by the engine; returns the device images and the oracle's view of them."""
    L, R = cs.synth_pair(seed, W, H, 2, 6, D - 1, 8)
    full = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1)
    disp = full.run(L, R)
    L4 = ((L.reshape(H // 4, 4, W // 4, 4).astype(np.int32).sum(axis=(1, 3)) + 8) >> 4).astype(np.uint8)
    R4 = ((R.reshape(H // 4, 4, W // 4, 4).astype(np.int32).sum(axis=(1, 3)) + 8) >> 4).astype(np.uint8)
    coarse = cs.Matcher(ctx, W // 4, H // 4, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7),
                        num_disp=max(8, (D // 4 + 7) // 8 * 8), tol=1)
    dc = coarse.run(L4, R4)
    imgs = (cs.DeviceImage.from_array(ctx, L4, cs.PIXEL_LUM_U8), cs.DeviceImage.from_array(ctx, R4, cs.PIXEL_LUM_U8),
            cs.DeviceImage.from_array(ctx, dc, cs.PIXEL_LUM_I16), cs.DeviceImage.from_array(ctx, disp, cs.PIXEL_LUM_I16))
    dfill = _fill(dc)
    xr = np.clip(np.arange(W // 4)[None, :] - dfill, 0, W // 4 - 1)
    inv = np.float32(1.0 / 255.0)  # the engine multiplies by the rounded reciprocal
    x0 = np.stack([L4.astype(np.float32) * inv, np.take_along_axis(R4, xr, axis=1).astype(np.float32) * inv], axis=-1)
    return imgs, x0.astype(np.float32), dfill.astype(np.float32), _labels(disp, W // 4, H // 4)


@pytest.mark.parametrize("conf_labels", [False, True])
def test_network_gradients_match_oracle(ctx, conf_labels):
    """Gradients of every layer against the oracle's network, with the confidence
    target from |e| <= tau or from given confidence labels (0, 1 and unset)."""
    W, H, D = 256, 128, 64
    net = cs.ProposalNet(ctx, W // 4, H // 4, max_frames=2)
    p = cs.cnn_init_params(1)
    p[-2:] = [0.3, -0.2]  # non-zero output biases
    net.set_params(p)
    xs, dfs, labs, cls = [], [], [], []
    rng = np.random.default_rng(5)
    for i, seed in enumerate((41, 42)):
        (dl, dr, dc, disp), x0, dfill, lab = _frame(ctx, seed, W, H, D)
        net.set_input(i, dl, dr, dc)
        net.set_labels(i, disp)
        cl = np.full(lab.shape, np.nan, np.float32)
        if conf_labels:
            conf = rng.integers(0, 3, lab.shape).astype(np.uint8)  # 2 = no target
            net.set_conf_labels(i, cs.DeviceImage.from_array(ctx, conf, cs.PIXEL_LUM_U8))
            cl = np.where(conf <= 1, conf.astype(np.float32), np.float32(np.nan))
        xs.append(x0)
        dfs.append(dfill)
        labs.append(lab)
        cls.append(cl)
    loss = net.train_step(2, 0.0)
    g = net.grads()
    x0, dfill, lab = np.stack(xs), np.stack(dfs), np.stack(labs)
    out, acts = cnn_ref.forward(p, x0)
    gout, ref_loss = cnn_ref.loss_grad(out, lab, dfill, np.stack(cls))
    ref = cnn_ref.backward(p, acts, gout)
    print(f"loss gpu {loss} oracle {ref_loss}", flush=True)
    assert int(loss[2]) == ref_loss["count"]
    assert abs(loss[0] - ref_loss["huber"]) <= 1e-3 * max(1.0, ref_loss["huber"])
    assert abs(loss[1] - ref_loss["bce"]) <= 1e-3 * max(1.0, ref_loss["bce"])
    k = 0
    for l, (ci, co, _) in enumerate(cs.CNN_LAYERS):
        n = co * 9 * ci + co
        a, b = g[k:k + n].astype(np.float64), ref[k:k + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  layer {l}: |grad| {np.linalg.norm(b):.3e} rel err {rel:.2e}", flush=True)
        assert rel <= 1e-3
        k += n
    assert k == cs.CNN_PARAMS


def test_training_reduces_loss(ctx):
    W, H, D = 256, 128, 64
    net = cs.ProposalNet(ctx, W // 4, H // 4, max_frames=4)
    net.set_params(cs.cnn_init_params(2))
    keep = []
    for i in range(4):
        imgs, _, _, _ = _frame(ctx, 100 + i, W, H, D)
        keep.append(imgs)
        net.set_input(i, *imgs[:3])
        net.set_labels(i, imgs[3])
    first = net.train_step(4, 2e-3)
    for _ in range(60):
        last = net.train_step(4, 2e-3)
    print(f"loss {first} -> {last}", flush=True)
    assert last[0] + last[1] < 0.8 * (first[0] + first[1])
    assert np.all(np.isfinite(net.params()))


def test_packed_bf16_conversion_rounds_like_the_integer_rule(ctx):
    """The one-launch network packs activations with the device's bf16 conversion
    instruction; the layer kernels and the CPU reference round with the integer
    nearest-even rule.  Every float bit pattern but NaN gives the same bf16."""
    n = C.c_uint64(123)
    assert cs.lib.stereo_bf16_rounding_check(ctx.handle, C.byref(n)) == 0
    assert n.value == 0


def test_fused_inference_equals_layer_launches(ctx):
    """stereo_cnn_forward runs the whole network in one launch (activations in LDS); its
    output equals five layer launches of stereo_conv3x3_bf16 on the same bf16 input and
    weights bit for bit, on frames whose size is not a multiple of the fused tile."""
    W, H, D = 536, 212, 64  # quarter resolution 134 x 53
    W4, H4 = W // 4, H // 4
    net = cs.ProposalNet(ctx, W4, H4, max_frames=2)
    p = cs.cnn_init_params(11)
    p[-2:] = [0.3, -0.2]
    net.set_params(p)
    xs = []
    for i, seed in enumerate((61, 62)):
        (dl, dr, dc, _), x0, _, _ = _frame(ctx, seed, W, H, D)
        net.set_input(i, dl, dr, dc)
        xs.append(x0)
    out = cs.DeviceBuffer(ctx, 2 * H4 * W4 * 2 * 4)
    net.forward_ptr(2, out.ptr)
    cs._check(cs.lib.stereo_synchronize(ctx.handle, None), ctx)
    got = out.download(np.float32, (2, H4, W4, 2))
    x = np.zeros((2, H4, W4, 16), np.float32)
    x[..., :2] = np.stack(xs)
    k = 0
    for l, (ci, co, dil) in enumerate(cs.CNN_LAYERS):
        w = np.zeros((16, 3, 3, 16), np.float32)
        w[:co, :, :, :ci] = p[k:k + co * 9 * ci].reshape(co, 3, 3, ci)
        k += co * 9 * ci
        b = np.zeros(16, np.float32)
        b[:co] = p[k:k + co]
        k += co
        _, _, x = _conv(ctx, x, w, b, dil, relu=l < 4)
    assert k == cs.CNN_PARAMS
    assert np.array_equal(got, x[..., :2]), int((got != x[..., :2]).sum())
    net.close()
