"""The batched-fitting workload (64 uneven instances, ds_multifit) against the
oracle.

1. A subset of the 64 instances that holds the largest and smallest images,
   the most and fewest splats, an instance without splats and one whose
   width and height are not multiples of the 16-pixel tile (about 3,400
   global tiles, so the tile ids take two 8-bit radix passes). After one
   forward + backward of all of them in one launch per phase, each instance
   is checked on its own: sorted ids and tile ranges bit-exact against the
   fp32 transcription of its Splat2D (tile_rasterize_layer.cu:52-162 per
   instance), its Splat2D against the FP64 projection, and the image,
   transmittance, last ids, dL/dC, dL/dSplat2D and parameter gradients
   against the FP64 oracle chain (compositing on the GPU's Splat2D) at
   |gpu - fp64| <= 1e-5 + 1e-4 |fp64| (each gradient field also to 2e-4 of
   its largest magnitude). Then one fused step (Adam inside the projection
   backward) from the same start against the FP64 Adam step.
2. All 64 instances for 20 graph-replayed steps: every value finite, final
   transmittance in [0, 1], every colour channel in [0, 1 - T], and every
   instance's loss lower than at the start.
"""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds
import oracle
from helpers import fp64_backward, fp64_forward_on, param_allowance
from test_gpu_bench_shape import check_close_allow, check_fields
from test_gpu_parity import MARGIN, check_close, report

pytestmark = pytest.mark.gpu

LR = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4, "lr_opacity": 6e-3}


def c3_subset():
    sizes = ds.batch_instances(64)
    area = [w * h for w, h, _, _ in sizes]
    count = [n for _, _, n, _ in sizes]
    pick = sorted({int(np.argmax(area)), int(np.argmin(area)), int(np.argmax(count)),
                   int(np.argmin(count))})
    sub = [sizes[i] for i in pick]
    sub.insert(1, (136, 100, 0, 3901))      # no splats
    sub.append((129, 97, 3000, 3902))        # ragged in both directions
    return sub


def make_multi(sizes, graph=False):
    inst = []
    for w, h, n, _ in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        inst.append((w, h, n, view, proj, cam))
    mf = ds.MultiFitter(inst, lr=LR, isect_cap=1 << 26)
    for k, (w, h, n, seed) in enumerate(sizes):
        mf.set_target(k, ds.target_image(seed, w, h, 32))
        mf.upload_params(k, ds.frustum_splats(seed, n, aspect=w / h))
    mf.set_tuning(1, 1)
    mf.set_graph(graph)
    return mf


def instance_pairs(st, k, tile_base, sizes):
    """Instance k's sorted ids (local splat indices) and tile ranges (local
    pair indices), cut out of the global pair list."""
    w, h, n, _ = sizes[k]
    nt = ((w + 15) // 16) * ((h + 15) // 16)
    ranges = st["tile_ranges"].reshape(-1, 2).astype(np.int64)
    lens = ranges[:, 1] - ranges[:, 0]
    tb = tile_base[k]
    pair_base = int(lens[:tb].sum())
    m = int(lens[tb:tb + nt].sum())
    r = ranges[tb:tb + nt]
    local = np.where(r[:, 1:2] > r[:, 0:1], r - pair_base, 0).astype(np.int32)
    return pair_base, m, local


def test_c3_instances_match_fp64_oracle():
    sizes = c3_subset()
    mf = make_multi(sizes)
    tile_base, pix_base, splat_base = mf.bases()
    w, h, _, _ = sizes[-1]
    n_tiles = tile_base[-1] + ((w + 15) // 16) * ((h + 15) // 16)
    assert n_tiles > 256  # the tile ids take two 8-bit radix passes
    mf.forward_backward()
    mf.synchronize()
    st = mf.raster_state()
    s2v, gs2v, grad_ptr = mf.views()
    n_all = sum(n for _, _, n, _ in sizes)
    s2_all = ds.download_ptr(s2v.pos_x, np.float32, 10 * n_all).reshape(10, n_all)
    gs2_all = ds.download_ptr(gs2v.pos_x, np.float32, 10 * n_all).reshape(10, n_all)
    npix_all = sum(w * h for w, h, _, _ in sizes)
    dldc_all = ds.download_ptr(grad_ptr, np.float32, 3 * npix_all)
    losses = mf.losses()
    total_pairs = 0
    grads0, params0, masks = [], [], [None] * len(sizes)
    for k, (w, h, n, seed) in enumerate(sizes):
        pb, sb = pix_base[k], splat_base[k]
        npix = w * h
        rgb = mf.image(k)
        T = st["T"][pb:pb + npix]
        pair_base, m, local = instance_pairs(st, k, tile_base, sizes)
        total_pairs += m
        params = ds.frustum_splats(seed, n, aspect=w / h)
        target = ds.target_image(seed, w, h, 32)
        g_gpu = mf.grads(k)
        grads0.append(g_gpu)
        params0.append(params)
        if n == 0:
            assert m == 0 and np.all(rgb == 0.0) and np.all(T == 1.0)
            assert abs(losses[k] - float(np.mean(target.astype(np.float64) ** 2) * 3)) <= \
                1e-4 * losses[k]
            continue
        s2 = np.ascontiguousarray(s2_all[:, sb:sb + n]).reshape(-1)
        ta = oracle.tile_assign_sort(s2, n, w, h)
        assert ta["n_isects"] == m, (k, ta["n_isects"], m)
        assert np.array_equal(st["flatten_ids"][pair_base:pair_base + m].astype(np.int64) - sb,
                              ta["ids"].astype(np.int64)), k
        assert np.array_equal(local.reshape(-1), ta["ranges"]), k
        view, proj, cam = ds.fitter_camera(w, h)
        pv = ds.mat_mul(proj, view).astype(np.float64)
        p64 = params.astype(np.float64)
        fw = fp64_forward_on(s2, p64, n, 0, cam.astype(np.float64), pv, w, h,
                             target.astype(np.float64), ta)
        live = np.repeat(fw["s2_fp64"][2 * n:3 * n] < 1e30, 10)
        check_close(f"c3_{w}x{h}_{n}_splat2d", s2, fw["s2_fp64"], live)
        bw = fp64_backward(fw, p64, n, 0, cam.astype(np.float64), pv, w, h,
                           amb_margin=MARGIN)
        ok_px = fw["comp"]["margin"] > MARGIN
        amb = bw["amb"]
        tag = f"c3_{w}x{h}_{n}"
        report(tag, n_isects=m, ambiguous_px=int((~ok_px).sum()),
               ambiguous_splats=int(amb.sum()), E_fwd=fw["comp"]["E"])
        assert (~ok_px).sum() <= max(3, npix // 250)
        check_close(f"{tag}_image", rgb, fw["comp"]["rgb"], np.repeat(ok_px, 3))
        check_close(f"{tag}_T", T, fw["comp"]["T"], ok_px)
        last = st["last_ids"][pb:pb + npix].astype(np.int64) - pair_base
        hit = ok_px & (fw["comp"]["T"] < 1.0)
        assert np.array_equal(last[hit], fw["comp"]["last"][hit].astype(np.int64)), k
        assert abs(losses[k] - fw["loss"]) <= 1e-5 + 1e-4 * fw["loss"], (k, losses[k])
        check_close(f"{tag}_dLdC", dldc_all[3 * pb:3 * (pb + npix)], fw["dLdC"],
                    np.repeat(ok_px, 3))
        # dL/dSplat2D and the parameter gradients with the decision-flip
        # allowance; the per-field check where the allowance is zero
        mask = np.ones(10 * n, bool)
        mask[2 * n:3 * n] = False
        gs2 = np.ascontiguousarray(gs2_all[:, sb:sb + n]).reshape(-1)
        check_close_allow(f"{tag}_dL_dsplat2d", gs2, bw["gs2"], bw["gs2_amb"], mask)
        smask = mask & np.tile(~amb, 10)
        check_fields(f"{tag}_dL_dsplat2d", np.where(smask, gs2, 0.0),
                     np.where(smask, bw["gs2"], 0.0), 10, n)
        p_allow = param_allowance(fw, p64, n, 0, cam.astype(np.float64), pv, bw["gs2_amb"])
        check_close_allow(f"{tag}_param_grads", g_gpu, bw["gparams"], p_allow,
                          np.ones(14 * n, bool))
        pmask = p_allow == 0
        check_fields(f"{tag}_param_grads", np.where(pmask, g_gpu, 0.0),
                     np.where(pmask, bw["gparams"], 0.0), 14, n)
        grads0[k] = bw["gparams"]
        masks[k] = pmask
    assert total_pairs == st["n_isects"]

    # one fused step from the same start: the first Adam step of every instance
    for k, (w, h, n, seed) in enumerate(sizes):
        mf.upload_params(k, params0[k])
    mf.step(1)
    mf.synchronize()
    groups = ((0, 3, LR["lr_pos"]), (3, 6, LR["lr_scale"]), (6, 10, LR["lr_rot"]),
              (10, 13, LR["lr_sh_dc"]), (13, 14, LR["lr_opacity"]))
    for k, (w, h, n, seed) in enumerate(sizes):
        if n == 0:
            continue
        p64 = params0[k].astype(np.float64)
        g64 = np.asarray(grads0[k], np.float64)
        p_ref = np.empty_like(p64)
        for f0, f1, lr in groups:
            sl = slice(f0 * n, f1 * n)
            z = np.zeros(sl.stop - sl.start)
            p_ref[sl], _, _ = oracle.adam(p64[sl], g64[sl], z, z, lr, 0.9, 0.999, 1e-8, 1)
        decided = (np.abs(g64) > 1e-6 * np.abs(g64).max()) & masks[k]
        check_close(f"c3_{w}x{h}_{n}_post_adam", mf.params(k), p_ref, decided)
    mf.close()


def test_c3_64_instances_20_steps_stay_valid_and_improve():
    sizes = ds.batch_instances(64)
    mf = make_multi(sizes, graph=True)
    mf.forward()
    l0 = mf.losses()
    mf.step(20)
    mf.synchronize()
    mf.forward()
    l1 = mf.losses()
    st = mf.raster_state()
    _, pix_base, _ = mf.bases()
    report("c3_64_instances_20_steps", loss_first=float(l0.mean()), loss_after=float(l1.mean()),
           n_isects=st["n_isects"])
    assert np.all(np.isfinite(l1)) and np.all(l1 < l0), (l0, l1)
    T = st["T"]
    assert np.all(np.isfinite(T)) and T.min() >= 0.0 and T.max() <= 1.0
    for k, (w, h, n, _) in enumerate(sizes):
        rgb = mf.image(k).reshape(-1, 3).astype(np.float64)
        t = T[pix_base[k]:pix_base[k] + w * h].astype(np.float64)
        assert np.all(np.isfinite(rgb)) and rgb.min() >= 0.0, k
        assert np.all(rgb <= (1.0 - t)[:, None] + 1e-5), k
        p = mf.params(k)
        assert np.all(np.isfinite(p)), k
    mf.close()
