"""Training of the learned initializer, and its use as the fitter's
starting point.

1. Pairs: `pairs` images of the synthetic-target generator (256x256, 32
   ellipses, seeds seed0 + i), each fitted for `fit_steps` plain-fitter steps
   (the fitter settings of the time-to-PSNR comparison) from the
   initializer's neutral splats (every head value 0: one splat per 4x4-pixel
   anchor), `batch` fits at a time in one ds_multifit.
2. Targets: each fitted slab through ds_initnet_head_targets, kept on the
   device with the images.
3. Supervised training: Adam on the mean squared error of the head values,
   one pair per step, `epochs` passes in a seeded shuffled order, the rate
   halving every `halve_every` steps.
4. Render-loss fine-tuning: `render_steps` steps of ds_initnet_train_step on
   the same images (the net's splats rendered and compared with the image,
   the gradient back through the renderer), the rate halving every
   `render_halve` steps.

`python -m diffsplat_amd.initnet_train --out W.npy` runs all four, measures
each stage's time to 30 dB on held-out seeds, and writes the weights (the
ds_initnet_get_params layout as bf16 bit patterns, the values the network's
kernels use) and W.json with the run's settings, timings and measurements.
`initialize` starts a fit from the trained initializer, with random splats
as the fallback; `time_to_psnr` times one fit to a PSNR.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time

import numpy as np

import diffsplat_amd as ds

W = H = 256
N_ELLIPSES = 32
EVAL_SEED0 = 90000
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "initnet_256.npy")
HV = ds.InitNet.HEAD_VALUES
FIT_LR = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4,
          "lr_opacity": 6e-3, "lr_pos_decay": 0.998}


def bf16_round(x):
    """float32 -> nearest bf16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def neutral_slab(width=W, height=H):
    """The initializer's splats for all-zero head values."""
    net = ds.InitNet(width, height, seed=0)
    net.set_params(np.zeros(net.num_params, np.float32))
    img = ds.DeviceArray.from_numpy(np.zeros(width * height * 3, np.float32))
    slab = ds.DeviceArray(4 * 14 * net.num_splats)
    net.infer(img, slab)
    ds.check(ds.lib().ds_synchronize(), "sync")
    out = slab.download(np.float32, 14 * net.num_splats)
    net.close()
    return out


def make_pairs(pairs, fit_steps, seed0=4000, batch=256, log=print):
    """Images (device, pairs x H*W*3) and head-value targets (device, pairs x
    splats*HV) of `pairs` plain fits from the initializer's neutral splats."""
    n = (W // 4) * (H // 4)
    start = neutral_slab()
    view, proj, cam = ds.fitter_camera(W, H)
    d_img = ds.DeviceArray(4 * pairs * W * H * 3)
    d_tgt = ds.DeviceArray(4 * pairs * n * HV)
    net = ds.InitNet(W, H, seed=0)
    d_slab = ds.DeviceArray(4 * 14 * n)
    psnr = []
    t0 = time.perf_counter()
    for b0 in range(0, pairs, batch):
        k = min(batch, pairs - b0)
        mf = ds.MultiFitter([(W, H, n, view, proj, cam)] * k, lr=FIT_LR, isect_cap=1 << 27)
        imgs = [ds.target_image(seed0 + b0 + i, W, H, N_ELLIPSES) for i in range(k)]
        for i in range(k):
            mf.set_target(i, imgs[i])
            mf.upload_params(i, start)
        mf.set_tuning(1, 1)
        mf.set_graph(True)
        mf.step(fit_steps)
        mf.forward()
        psnr += [ds.psnr_from_loss(float(x)) for x in mf.losses()]
        for i in range(k):
            j = b0 + i
            ds.check(ds.lib().ds_memcpy_htod(d_img.ptr + 4 * j * W * H * 3,
                                             imgs[i].ctypes.data, 4 * W * H * 3), "upload")
            d_slab.upload(mf.params(i))
            net.head_targets(d_slab, d_tgt.ptr + 4 * j * n * HV)
        mf.close()
        log(f"[initnet] pairs {b0 + k}/{pairs}: mean fitted PSNR {np.mean(psnr):.2f} dB, "
            f"{time.perf_counter() - t0:.1f} s")
    ds.check(ds.lib().ds_synchronize(), "sync")
    net.close()
    return d_img, d_tgt, {"pairs": pairs, "fit_steps": fit_steps, "seed0": seed0,
                          "fitted_psnr_db_mean": round(float(np.mean(psnr)), 3),
                          "seconds": round(time.perf_counter() - t0, 2)}


def train(net, d_img, d_tgt, pairs, epochs=3, lr=2e-3, halve_every=10000, seed=0, log=print):
    """Supervised Adam steps over the pairs; returns the mean loss per epoch."""
    n = net.num_splats
    rng = np.random.default_rng(seed)
    d_loss = ds.DeviceArray(4)
    losses = []
    step = 0
    t0 = time.perf_counter()
    for e in range(epochs):
        d_loss.zero()
        for j in rng.permutation(pairs):
            net.train_step_supervised(d_img.ptr + 4 * int(j) * W * H * 3,
                                      d_tgt.ptr + 4 * int(j) * n * HV,
                                      lr * 0.5 ** (step / halve_every), d_loss)
            step += 1
        ds.check(ds.lib().ds_synchronize(), "sync")
        losses.append(float(d_loss.download(np.float32, 1)[0]) / pairs)
        log(f"[initnet] epoch {e + 1}/{epochs}: mean head loss {losses[-1]:.4f}, "
            f"{time.perf_counter() - t0:.1f} s")
    return losses, round(time.perf_counter() - t0, 2)


def save_weights(path, params):
    """Weights as bf16 bit patterns (uint16 .npy): the values the network's
    kernels use."""
    bits = (bf16_round(params).view(np.uint32) >> 16).astype(np.uint16)
    np.save(path, bits, allow_pickle=False)


def load(path, width=W, height=H):
    """An InitNet with the weights of a saved training run."""
    bits = np.load(path, allow_pickle=False).astype(np.uint32) << 16
    net = ds.InitNet(width, height, seed=0)
    net.set_params(bits.view(np.float32))
    return net


def render_finetune(net, d_img, count, steps, lr=2e-3, halve_every=9000, log=print):
    """Render-loss steps (ds_initnet_train_step) cycling over `count` device
    images; returns the seconds taken."""
    n = net.num_splats
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam)
    t0 = time.perf_counter()
    for k in range(steps):
        img = C.c_void_p(d_img.ptr + 4 * (k % count) * W * H * 3)
        ds.check(ds.lib().ds_initnet_train_step(net.h, fit.h, img,
                                                lr * 0.5 ** (k / halve_every)),
                 "ds_initnet_train_step")
        if (k + 1) % 10000 == 0:
            fit.synchronize()
            log(f"[initnet] render steps {k + 1}/{steps}, {time.perf_counter() - t0:.1f} s")
    fit.synchronize()
    fit.close()
    return round(time.perf_counter() - t0, 2)


def initialize(fit, net, d_image, random_slab):
    """Starts `fit` from the initializer's splats for the image, with the
    random init as the fallback: both are rendered once and the one with the
    lower loss is kept. Returns True when the initializer's splats were kept.
    The fitter's optimizer starts from zero either way."""
    fit.upload_params(random_slab)
    fit.forward()
    random_loss = fit.loss()
    pslab, _ = fit.param_slabs()
    net.infer(d_image, pslab, fit.stream())
    fit.reset_optimizer()
    fit.forward()
    if fit.loss() <= random_loss:
        return True
    fit.upload_params(random_slab)
    return False


def time_to_psnr(image, seed, net=None, d_image=None, target_db=30.0, max_steps=3000,
                 check_every=10):
    """Wall time for one fit of `image` (W x H) to reach target_db PSNR, from
    the random frustum splats of `seed` or, with `net`, from `initialize`
    (inference and the fallback's two renders inside the timed region). HIP
    graph on (its capture on the first step is timed too); the loss of the
    last step's forward is read every check_every steps (one host sync).
    Returns (steps, seconds, final PSNR, initial PSNR, initializer kept)."""
    n = (W // 4) * (H // 4)
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam, lr=FIT_LR)
    fit.set_target(image)
    rnd = ds.frustum_splats(seed, n, aspect=W / H)
    fit.upload_params(rnd)
    fit.set_graph(True)
    fit.synchronize()
    t0 = time.perf_counter()
    kept = initialize(fit, net, d_image, rnd) if net is not None else False
    fit.forward()
    p0 = ds.psnr_from_loss(fit.loss())
    steps, psnr = 0, p0
    while steps < max_steps and psnr < target_db:
        fit.step(check_every)
        steps += check_every
        psnr = ds.psnr_from_loss(fit.loss())
    el = time.perf_counter() - t0
    fit.close()
    return steps, el, psnr, p0, kept


def evaluate(net, count, seed0=EVAL_SEED0, target_db=30.0):
    """Mean time / steps to target_db over `count` held-out seeds, from the
    initializer (net) or from random splats (net None)."""
    r = []
    for i in range(count):
        img = ds.target_image(seed0 + i, W, H, N_ELLIPSES)
        d = ds.DeviceArray.from_numpy(img) if net is not None else None
        r.append(time_to_psnr(img, seed0 + i, net, d, target_db))
    a = np.array([x[:4] for x in r], np.float64)
    return {"mean_seconds": round(float(a[:, 1].mean()), 5),
            "mean_steps": round(float(a[:, 0].mean()), 1),
            "reached": int((a[:, 2] >= target_db).sum()), "fits": count,
            "initial_psnr_db": round(float(a[:, 3].mean()), 3),
            "initializer_kept": int(sum(x[4] for x in r))}


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=WEIGHTS, help="weights file (.npy); settings go to .json")
    ap.add_argument("--pairs", type=int, default=10000)
    ap.add_argument("--fit-steps", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--halve-every", type=int, default=30000)
    ap.add_argument("--render-steps", type=int, default=30000)
    ap.add_argument("--render-lr", type=float, default=2e-3)
    ap.add_argument("--render-halve", type=int, default=9000)
    ap.add_argument("--eval", type=int, default=64, help="held-out seeds per measurement")
    ap.add_argument("--compare", action="store_true",
                    help="also train a net by render loss alone and measure it")
    a = ap.parse_args()
    meas = {"plain": evaluate(None, a.eval)}
    d_img, d_tgt, data = make_pairs(a.pairs, a.fit_steps, batch=a.batch)
    net = ds.InitNet(W, H, seed=1)
    losses, train_s = train(net, d_img, d_tgt, a.pairs, a.epochs, a.lr, a.halve_every)
    del d_tgt
    meas["supervised"] = evaluate(net, a.eval)
    print(f"[initnet] supervised: {meas['supervised']}", flush=True)
    render_s = render_finetune(net, d_img, a.pairs, a.render_steps, a.render_lr, a.render_halve)
    meas["supervised+render"] = evaluate(net, a.eval)
    print(f"[initnet] supervised+render: {meas['supervised+render']}", flush=True)
    p = net.params()
    save_weights(a.out, p)
    if a.compare:
        alone = ds.InitNet(W, H, seed=1)
        alone_s = render_finetune(alone, d_img, a.pairs, a.render_steps, a.render_lr,
                                  a.render_halve)
        meas["render_only"] = evaluate(alone, a.eval)
        meas["render_only"]["train_seconds"] = alone_s
        alone.close()
    meta = {"data": data, "epochs": a.epochs, "lr": a.lr, "halve_every": a.halve_every,
            "epoch_losses": [round(x, 5) for x in losses], "train_seconds": train_s,
            "render_steps": a.render_steps, "render_lr": a.render_lr,
            "render_halve": a.render_halve, "render_seconds": render_s,
            "params": int(p.size), "image": [W, H], "ellipses": N_ELLIPSES,
            "fit_lr": FIT_LR, "eval_seed0": EVAL_SEED0, "time_to_30db": meas}
    with open(a.out.rsplit(".", 1)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
