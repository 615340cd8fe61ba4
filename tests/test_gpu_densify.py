"""Adaptive density control on the GPU (ds_fitter_densify) against the oracle.

The fitter accumulates per-splat gradient statistics inside its captured
step, then one densification classifies, scans and scatters on the device.
Checked against the oracle on the GPU's own inputs: the statistics of one
step (|dL/d(NDC centre)| of every splat with a tile, one visit), the actions
and new count bit-exact, survivors and clones bit-exact with their Adam
moments, split samples within 1e-5 + 1e-4 |x| of the FP64 draw; then the fit
continues from the new splats with the step counter kept and its loss
falling."""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds
import oracle

pytestmark = pytest.mark.gpu

W = H = 128
N = 2000
LR = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4, "lr_opacity": 6e-3}


def make_fitter(graph=True):
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, N, view=view, proj=proj, cam_pos=cam, lr=LR)
    fit.set_target(ds.target_image(11, W, H, 16))
    params = ds.frustum_splats(11, N, aspect=1.0, ls_mean=-2.6).reshape(14, N).copy()
    params[13, :100] = -7.0  # faint: pruned
    fit.upload_params(params.reshape(-1))
    fit.set_graph(graph)
    return fit, view, proj, cam


def test_one_step_statistics_are_the_centre_gradient_lengths():
    fit, *_ = make_fitter()
    fit.set_densify_stats(True)
    fit.step(1)
    fit.synchronize()
    acc, vis = fit.densify_stats()
    from test_gpu_bench_shape import fitter_views
    s2v, gs2v, _ = fitter_views(fit)
    s2 = ds.download_ptr(s2v.pos_x, np.float32, 10 * N)
    gs2 = ds.download_ptr(gs2v.pos_x, np.float32, 10 * N).reshape(10, N)
    ta = oracle.tile_assign_sort(s2, N, W, H)
    ids_with_tiles = np.zeros(N, bool)
    ids_with_tiles[np.unique(ta["ids"])] = True
    assert np.array_equal(vis, ids_with_tiles.astype(np.uint32))
    want = np.where(ids_with_tiles, np.sqrt(gs2[0].astype(np.float64) ** 2 +
                                            gs2[1].astype(np.float64) ** 2), 0.0)
    np.testing.assert_allclose(acc, want, rtol=1e-6, atol=1e-12)
    fit.close()


def test_densify_matches_the_oracle_and_the_fit_continues():
    fit, view, proj, cam = make_fitter()
    fit.set_densify_stats(True)
    fit.step(40)
    fit.synchronize()
    acc, vis = fit.densify_stats()
    p0 = fit.params()
    m0, v0 = fit.moments()
    t0 = fit.step_count()
    g = np.where(vis > 0, acc / np.maximum(vis, 1), 0)
    thr = float(np.quantile(g[vis > 0], 0.7))  # about 30% of the visible splats densify
    ls = p0.reshape(14, N)[3:6].max(axis=0)
    extent = 4.0
    pd = float(np.exp(np.median(ls))) / extent  # clone the smaller half, split the rest
    cfg = dict(grad_threshold=thr, percent_dense=pd, scene_extent=extent, min_opacity=0.005,
               max_scale_frac=0.0, split_scale_div=1.6, seed=99)
    res, acts = fit.densify(**cfg)
    p = oracle.densify_params(**cfg)
    ref = oracle.densify(p0, m0, v0, N, 0, acc, vis, p)
    print(f"[parity] densify: {res}")
    assert np.array_equal(acts, ref["action"])
    assert res["count_after"] == ref["n_out"] == fit.count
    assert res["pruned"] >= 100 and res["cloned"] > 0 and res["split"] > 0
    n1 = ref["n_out"]
    p1 = fit.params().reshape(14, n1).astype(np.float64)
    m1, v1 = fit.moments()
    R = ref["params"].reshape(14, n1)
    split_cols = np.zeros(n1, bool)
    nkeep = int(((ref["action"] == 1) | (ref["action"] == 2)).sum())
    src_new = np.sort(np.concatenate([np.where(ref["action"] == 2)[0],
                                      np.repeat(np.where(ref["action"] == 3)[0], 2)]),
                      kind="stable")
    split_cols[nkeep:] = ref["action"][src_new] == 3
    exact = ~split_cols
    assert np.array_equal(p1[:, exact], R[:, exact])
    assert np.array_equal(m1, ref["m"]) and np.array_equal(v1, ref["v"])
    d = np.abs(p1[:, split_cols] - R[:, split_cols])
    assert np.all(d <= 1e-5 + 1e-4 * np.abs(R[:, split_cols])), float(d.max())
    print(f"[parity] densify_split_samples: max_abs={d.max():.3e} n={int(split_cols.sum())}")
    # the fit goes on at the new count: step counter kept, graph recaptured
    assert fit.step_count() == t0
    fit.forward()
    l0 = fit.loss()
    fit.step(60)
    fit.synchronize()
    assert fit.step_count() == t0 + 60
    fit.forward()
    l1 = fit.loss()
    assert np.isfinite(l1) and l1 < l0, (l0, l1)
    fit.close()
