"""Trains the learned disparity-proposal network on the GPU with the engine's own
kernels (BF16 MFMA forward and backward, Adam) and writes its parameters.

Data: seeded synthetic pairs (the engine's generator, constant and slanted
regions) at 960x540.  Network input: the pair at 1/4 resolution, the right image
warped by the census match at that resolution (D=32).  Targets, per quarter pixel:
the centre residual against the full-range census matcher's checked disparity map
(9x7 census, 9x9 window, D=128; Huber), and the confidence that the band search
around the quarter-resolution match gives the full-range result there (all 16
pixels equal; BCE).  At run time a tile whose smallest confidence logit is below
the calibrated threshold searches the full range.

    python module.clutec.cuda_amd/train_proposal.py --frames 10000 --steps 3000 \
        --out module.clutec.cuda_amd/weights/proposal_cnn.bin

Writes <out> (STEREO_CNN_PARAMS float32, little endian) and <out>.json (settings,
loss and the calibration of the threshold at 1920x1080, D=256).
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import clutec_stereo as cs  # noqa: E402


def downsample4(img):
    H, W = img.shape
    return ((img.reshape(H // 4, 4, W // 4, 4).astype(np.int32).sum(axis=(1, 3)) + 8) >> 4).astype(np.uint8)


class FramePool:
    """Prepared training frames resident in device memory: quarter-resolution L4,
    R4, the quarter-resolution census match and the full-resolution label map."""

    def __init__(self, ctx, W, H, D, frames, seed0, threads=16, log=print):
        self.ctx, self.W, self.H = ctx, W, H
        Dc = max(8, (D // 4 + 7) // 8 * 8)
        full = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1)
        coarse = cs.Matcher(ctx, W // 4, H // 4, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=Dc,
                            tol=1)
        dl = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U8)
        dr = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_U8)
        # the band search around the quarter-resolution match with no fallback (the
        # network's output does not enter: bands come from the coarse map)
        band = cs.LearnedMatcher(ctx, W, H, cs.cnn_init_params(0), num_disp=D, band=8, chunk=32, logit_thresh=-1e30)
        bout = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
        self.agree = 0.0
        self.items = []
        t0 = time.time()

        def make(seed):
            kind = 1 if seed % 2 else 0  # slanted and constant regions
            return cs.synth_pair(seed, W, H, kind, 12, D - 1, 8)

        def pairs():
            # generate in chunks so that at most a few hundred host pairs are held
            with ThreadPoolExecutor(max_workers=threads) as ex:
                for c in range(seed0, seed0 + frames, 256):
                    yield from ex.map(make, range(c, min(seed0 + frames, c + 256)))

        for i, (L, R) in enumerate(pairs()):
            dl.upload(L)
            dr.upload(R)
            disp = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
            full.process(disp, None, None, dl, dr)
            L4 = cs.DeviceImage.from_array(ctx, downsample4(L), cs.PIXEL_LUM_U8)
            R4 = cs.DeviceImage.from_array(ctx, downsample4(R), cs.PIXEL_LUM_U8)
            dc = cs.DeviceImage(ctx, W // 4, H // 4, cs.PIXEL_LUM_I16)
            coarse.process(dc, None, None, L4, R4)
            band.process(bout, None, None, dl, dr)
            same = (bout.download() == disp.download()).reshape(H // 4, 4, W // 4, 4).all(axis=(1, 3))
            self.agree += float(same.mean()) / frames
            conf = cs.DeviceImage.from_array(ctx, same.astype(np.uint8), cs.PIXEL_LUM_U8)
            self.items.append((L4, R4, dc, disp, conf))
            if (i + 1) % 1000 == 0:
                log(f"  prepared {i + 1}/{frames} frames ({time.time() - t0:.1f}s)")
        ctx.synchronize()


def train(ctx, pool, steps, batch, lr, seed, log=print, init=None):
    W4, H4 = pool.W // 4, pool.H // 4
    net = cs.ProposalNet(ctx, W4, H4, max_frames=batch)
    net.set_params(cs.cnn_init_params(seed) if init is None else init)
    rng = np.random.default_rng(seed)
    history = []
    t0 = time.time()
    for step in range(steps):
        idx = rng.choice(len(pool.items), size=batch, replace=False)
        for slot, k in enumerate(idx):
            L4, R4, dc, disp, conf = pool.items[k]
            net.set_input(slot, L4, R4, dc)
            net.set_labels(slot, disp)
            net.set_conf_labels(slot, conf)
        # cosine decay to 5 % of the initial rate
        rate = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * step / max(1, steps - 1))))
        loss = net.train_step(batch, float(rate))
        history.append([float(v) for v in loss])
        if step % 200 == 0 or step == steps - 1:
            log(f"  step {step:5d} lr {rate:.2e} huber {loss[0]:.4f} bce {loss[1]:.4f} "
                f"target-1 fraction {loss[3]:.4f} ({time.time() - t0:.1f}s)")
    return net.params(), history


def calibrate(ctx, params, W, H, D, seeds, thresholds, min_agree=0.995, band=8, log=print):
    """Agreement with the full-range matcher and searched fraction of the cost volume
    per confidence threshold, on held-out frames at the deployment size (bands around
    the quarter-resolution census map).  After the given grid, the interval between
    the largest failing and the smallest passing threshold is scanned in tenths.
    Returns the smallest threshold (most tiles on their band) at which every frame
    reaches min_agree, and the table."""
    full = cs.Matcher(ctx, W, H, algorithm=cs.ALGO_CENSUS, window=(9, 9), census=(9, 7), num_disp=D, tol=1)
    frames = []
    for seed in seeds:
        L, R = cs.synth_pair(seed, W, H, seed % 3, 24, D - 1, 8)
        dl = cs.DeviceImage.from_array(ctx, L, cs.PIXEL_LUM_U8)
        dr = cs.DeviceImage.from_array(ctx, R, cs.PIXEL_LUM_U8)
        ref = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)
        full.process(ref, None, None, dl, dr)
        frames.append((dl, dr, ref.download()))
    out = cs.DeviceImage(ctx, W, H, cs.PIXEL_LUM_I16)

    def row_for(t):
        lm = cs.LearnedMatcher(ctx, W, H, params, num_disp=D, band=band, chunk=32, logit_thresh=t)
        agree, frac = [], []
        for dl, dr, ref in frames:
            lm.process(out, None, None, dl, dr)
            agree.append(float((out.download() == ref).mean()))
            frac.append(lm.stats()["evals"] / float(W * H * D))
        row = {"thresh": round(float(t), 3), "agree_mean": round(float(np.mean(agree)), 5),
               "agree_min": round(float(np.min(agree)), 5), "evals_frac": round(float(np.mean(frac)), 4)}
        log(f"  threshold {t:+.2f}: agreement {row['agree_mean']:.5f} (min {row['agree_min']:.5f}) "
            f"searched {row['evals_frac']:.3f}")
        return row

    table = [row_for(t) for t in thresholds]
    ok = [r["thresh"] for r in table if r["agree_min"] >= min_agree]
    bad = [r["thresh"] for r in table if r["agree_min"] < min_agree]
    if ok and bad and max(bad) < min(ok):
        lo, hi = max(bad), min(ok)
        table += [row_for(lo + (hi - lo) * k / 10.0) for k in range(1, 10)]
    table.sort(key=lambda r: r["thresh"])
    ok = [r for r in table if r["agree_min"] >= min_agree]
    best = min(ok, key=lambda r: r["thresh"]) if ok else max(table, key=lambda r: r["thresh"])
    return best["thresh"], table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=10000)
    ap.add_argument("--width", type=int, default=960)
    ap.add_argument("--height", type=int, default=540)
    ap.add_argument("--disparities", type=int, default=128)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--seed0", type=int, default=10000, help="first synthetic pair seed")
    ap.add_argument("--out", required=True)
    ap.add_argument("--calib-frames", type=int, default=8, help="held-out 1920x1080 D=256 frames for the threshold")
    ap.add_argument("--band", type=int, default=8, help="band of the searched disparities around each centre")
    ap.add_argument("--calibrate-only", action="store_true",
                    help="keep the parameters in --out and only recalibrate the threshold")
    a = ap.parse_args()
    ctx = cs.Context(0)
    log = lambda s: print(s, flush=True)
    thresholds = [float(t) for t in np.arange(-6.0, 8.01, 0.5)]
    if a.calibrate_only:
        params = np.fromfile(a.out, "<f4")
        with open(a.out + ".json") as f:
            meta = json.load(f)
        log(f"calibrating the confidence threshold at 1920x1080, D=256, band {a.band}")
        thr, table = calibrate(ctx, params, 1920, 1080, 256, range(20000, 20000 + a.calib_frames), thresholds,
                               band=a.band, log=log)
        meta.update({"logit_thresh": thr, "calibration": table, "band": a.band, "centre_residual": 0,
                     "calibration_rule": "smallest threshold with every held-out frame at agreement >= 0.995"})
        with open(a.out + ".json", "w") as f:
            json.dump(meta, f, indent=1)
        log(json.dumps({"logit_thresh": thr}))
        return
    log(f"preparing {a.frames} frames {a.width}x{a.height} D={a.disparities}")
    t0 = time.time()
    pool = FramePool(ctx, a.width, a.height, a.disparities, a.frames, a.seed0, log=log)
    t_data = time.time() - t0
    log(f"quarter pixels where the coarse band agrees with the full range: {pool.agree:.4f}")
    log(f"training {a.steps} steps, batch {a.batch}")
    t1 = time.time()
    params, hist = train(ctx, pool, a.steps, a.batch, a.lr, a.seed, log=log)
    t_train = time.time() - t1
    log("calibrating the confidence threshold at 1920x1080, D=256")
    thr, table = calibrate(ctx, params, 1920, 1080, 256, range(20000, 20000 + a.calib_frames), thresholds,
                           band=a.band, log=log)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    params.astype("<f4").tofile(a.out)
    tail = np.array(hist[-100:])
    meta = {
        "params": int(params.size), "frames": a.frames, "size": [a.width, a.height], "disparities": a.disparities,
        "steps": a.steps, "batch": a.batch, "lr": a.lr, "seed": a.seed, "seed0": a.seed0,
        "first_loss": hist[0], "final_loss_mean100": tail.mean(axis=0).tolist(),
        "data_seconds": round(t_data, 1), "train_seconds": round(t_train, 1), "band_agree_fraction": round(pool.agree, 5),
        "confidence_target": "band search around the quarter-resolution match equals the full range (4x4 block)",
        "logit_thresh": thr, "calibration": table, "band": a.band, "centre_residual": 0,
        "calibration_rule": "smallest threshold with every held-out frame at agreement >= 0.995",
        "loss_fields": ["huber", "bce", "labelled_pixels", "confident_target_fraction"],
    }
    with open(a.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))


if __name__ == "__main__":
    main()
