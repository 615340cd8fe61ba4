"""Learned initializer (ds_initnet) on the GPU against the numpy reference in
tests/initnet_ref.py, which rounds to bf16 where the kernels store bf16 (so
only the fp32 accumulation order differs, plus the rare activation that then
rounds to the neighbouring bf16 value): head values, the splat slab, every
weight and bias gradient, and a short training run through the renderer."""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds
import initnet_ref as R

pytestmark = pytest.mark.gpu


def random_params(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout, ks, _ in R.LAYERS:
        k = ks * ks * cin
        parts += [rng.normal(0, np.sqrt(2.0 / k), cout * k), rng.normal(0, 0.05, cout)]
    return np.concatenate(parts).astype(np.float32)


def setup(W, H, seed):
    net = ds.InitNet(W, H, seed=seed)
    p = random_params(seed)
    net.set_params(p)
    img = ds.target_image(seed + 7, W, H, 8)
    return net, p, img, ds.DeviceArray.from_numpy(img)


@pytest.mark.parametrize("W,H", [(64, 64), (96, 48)])
def test_forward_matches_reference(W, H):
    net, p, img, d_img = setup(W, H, 3)
    n = net.num_splats
    assert n == (W // 4) * (H // 4)
    d_slab = ds.DeviceArray(4 * 14 * n)
    net.infer(d_img, d_slab)
    h = net.head().reshape(n, R.HEAD)
    h_ref, _ = R.forward(p, img, W, H)
    err = float(np.abs(h - h_ref).max())
    scale = float(np.abs(h_ref).max())
    print(f"[parity] initnet_head_{W}x{H}: max_abs={err:.3e} scale={scale:.3e}")
    assert scale > 0.1 and err <= 1e-2 * scale
    slab = d_slab.download(np.float32, 14 * n)
    ref = R.head_fwd(h.astype(np.float64), R.geometry(W, H))
    np.testing.assert_allclose(slab, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("W,H", [(64, 64), (96, 48)])
def test_gradients_match_reference(W, H):
    net, p, img, d_img = setup(W, H, 5)
    n = net.num_splats
    gs = np.random.default_rng(9).normal(0, 1e-2, 14 * n).astype(np.float32)
    g = net.grads(d_img, ds.DeviceArray.from_numpy(gs))
    h_ref, cache = R.forward(p, img, W, H)
    g_ref = R.backward(cache, h_ref, R.geometry(W, H), gs)
    o = 0
    for i, (cin, cout, ks, _) in enumerate(R.LAYERS):
        for part, size in (("w", cout * ks * ks * cin), ("b", cout)):
            a, b = g[o:o + size], g_ref[o:o + size]
            rel = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))
            print(f"[parity] initnet_grad_{W}x{H}_layer{i}_{part}: rel_l2={rel:.3e}")
            assert rel <= 2e-2, (i, part, rel)
            o += size
    assert o == g.size


def test_training_through_the_renderer_reduces_loss():
    W = H = 64
    net = ds.InitNet(W, H, seed=1)
    n = net.num_splats
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam)
    imgs = [ds.DeviceArray.from_numpy(ds.target_image(100 + i, W, H, 8)) for i in range(8)]
    pslab, _ = fit.param_slabs()

    def mean_loss():
        total = 0.0
        for d in imgs:
            net.infer(d, pslab, fit.stream())
            fit.set_target_device(d.ptr)
            fit.forward()
            total += fit.loss()
        return total / len(imgs)

    before = mean_loss()
    for k in range(400):
        net.train_step(fit, imgs[k % len(imgs)], 2e-3)
    fit.synchronize()
    after = mean_loss()
    print(f"[parity] initnet_train: loss_before={before:.5f} after={after:.5f}")
    assert np.isfinite(after) and after < 0.7 * before
    fit.close()
    net.close()
