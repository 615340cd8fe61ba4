"""Many fits as one (multi-instance rasterizer, ds_activ_project_*_multi,
ds_multifit) on the GPU. Per instance every result must equal the
one-instance path on the same splats: the activation/projection, sorted ids,
tile ranges, images, transmittance and last ids bit for bit (compositing of
a tile runs the same code on the same list), gradients and fitted parameters
up to the order of float atomics."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import diffsplat_amd as ds
from helpers import close

pytestmark = pytest.mark.gpu

LR = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4, "lr_opacity": 6e-3}
# (width, height, splats, seed): ragged sizes, one instance without splats,
# one whose width is not a multiple of the tile
SIZES = [(96, 64, 700, 21), (40, 72, 0, 22), (150, 90, 1800, 23), (64, 64, 300, 24)]


def cam_table(sizes):
    tab = []
    for w, h, _, _ in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        tab += list(ds.mat_mul(proj, view)) + list(cam) + [0.0]
    return np.array(tab, np.float32)


def combined_slab(slabs, counts):
    """Instance slabs (rows of count floats) -> one slab with rows of sum(counts)."""
    rows = 14
    return np.concatenate([s.reshape(rows, n) for s, n in zip(slabs, counts)],
                          axis=1).reshape(-1)


def single_stage(w, h, n, seed):
    """One-instance activation/projection + rasterizer forward/backward."""
    params = ds.frustum_splats(seed, n, aspect=w / h)
    target = ds.target_image(seed, w, h, 12)
    view, proj, cam = ds.fitter_camera(w, h)
    pv = ds.mat_mul(proj, view)
    out = {"params": params, "target": target}
    if n == 0:
        return out
    d_p = ds.DeviceArray.from_numpy(params)
    d_s2 = ds.DeviceArray(4 * 10 * n)
    d_gs2 = ds.DeviceArray(4 * 10 * n)
    d_gs2.zero()
    d_rgb = ds.DeviceArray(4 * 3 * w * h)
    d_tgt = ds.DeviceArray.from_numpy(target)
    d_loss = ds.DeviceArray(4)
    d_grad = ds.DeviceArray(4 * 3 * w * h)
    g = ds.gauss_view(d_p, n, 0)
    s2 = ds.splat2d_view(d_s2, n)
    gs2 = ds.splat2d_view(d_gs2, n)
    ds.check(ds.lib().ds_activ_project_fwd(g, s2, 0, ds.fptr(cam), ds.fptr(pv), None), "fwd")
    ras = ds.Rasterizer(w, h, n, isect_cap=1 << 20, keep_keys=True)
    ras.forward(s2, d_rgb, d_tgt, d_loss, d_grad)
    ras.backward(s2, d_grad, gs2)
    st = ras.state()
    out.update(st)
    out["s2"] = d_s2.download(np.float32, 10 * n)
    out["rgb"] = d_rgb.download(np.float32, 3 * w * h)
    out["loss"] = float(d_loss.download(np.float32, 1)[0])
    out["gs2"] = d_gs2.download(np.float32, 10 * n)
    ras.close()
    return out


def test_multi_raster_matches_one_instance_rasterizers():
    singles = [single_stage(*s) for s in SIZES]
    counts = [s[2] for s in SIZES]
    total = sum(counts)
    mr = ds.MultiRasterizer([(w, h, n) for w, h, n, _ in SIZES], isect_cap=1 << 21,
                            keep_keys=True)
    ds.check(ds.lib().ds_raster_set_tuning(mr.h, 2, 4), "set_tuning")
    # activation + projection of all instances in one launch
    slab = combined_slab([s["params"] for s in singles], counts)
    d_p = ds.DeviceArray.from_numpy(slab)
    d_s2 = ds.DeviceArray(4 * 10 * total)
    d_gs2 = ds.DeviceArray(4 * 10 * total)
    d_gs2.zero()
    d_cam = ds.DeviceArray.from_numpy(cam_table(SIZES))
    g = ds.gauss_view(d_p, total, 0)
    s2 = ds.splat2d_view(d_s2, total)
    gs2 = ds.splat2d_view(d_gs2, total)
    ds.check(ds.lib().ds_activ_project_fwd_multi(g, s2, 0, d_cam.ptr, mr.splat_instances(),
                                                 None), "ds_activ_project_fwd_multi")
    s2_all = d_s2.download(np.float32, 10 * total).reshape(10, total)
    for (w, h, n, _), s, b in zip(SIZES, singles, mr.splat_base):
        if n:
            assert np.array_equal(s2_all[:, b:b + n], s["s2"].reshape(10, n)), (w, h, n)
    targets = np.concatenate([s["target"] for s in singles])
    d_rgb = ds.DeviceArray(4 * 3 * mr.num_pixels)
    d_tgt = ds.DeviceArray.from_numpy(targets)
    d_loss = ds.DeviceArray(4 * len(SIZES))
    d_grad = ds.DeviceArray(4 * 3 * mr.num_pixels)
    mr.forward(s2, d_rgb, d_tgt, d_loss, d_grad)
    mr.backward(s2, d_grad, gs2)
    st = mr.state()
    rgb = d_rgb.download(np.float32, 3 * mr.num_pixels)
    losses = d_loss.download(np.float32, len(SIZES))
    gs2_all = d_gs2.download(np.float32, 10 * total).reshape(10, total)
    assert st["n_isects"] == sum(s.get("n_isects", 0) for s in singles)
    # heavy-first dispatch over the tiles of all instances
    counts = st["tile_ranges"].reshape(-1, 2) @ np.array([-1, 1])
    order = st["tile_order"]
    assert np.array_equal(np.sort(order), np.arange(mr.num_tiles))
    bucket = np.where(counts > 0, np.floor(np.log2(np.maximum(counts, 1))) + 1, 0)
    assert np.all(np.diff(bucket[order]) <= 0)
    ids, ranges, keys = st["flatten_ids"], st["tile_ranges"].reshape(-1, 2), st["keys"]
    pair_base = 0
    for k, ((w, h, n, _), s) in enumerate(zip(SIZES, singles)):
        tb, pb, sb = mr.tile_base[k], mr.pix_base[k], mr.splat_base[k]
        nt = ((w + 15) // 16) * ((h + 15) // 16)
        npix = w * h
        if n == 0:
            assert np.all(ranges[tb:tb + nt] == 0)
            assert np.all(rgb[3 * pb:3 * (pb + npix)] == 0) and losses[k] > 0
            continue
        m = s["n_isects"]
        assert np.array_equal(ids[pair_base:pair_base + m], s["flatten_ids"] + sb), k
        tile_of = (keys[pair_base:pair_base + m] >> np.uint64(32)).astype(np.int64)
        assert np.array_equal(tile_of - tb, (s["keys"] >> np.uint64(32)).astype(np.int64))
        assert np.array_equal(keys[pair_base:pair_base + m] & np.uint64(0xFFFFFFFF),
                              s["keys"] & np.uint64(0xFFFFFFFF))
        r1 = s["tile_ranges"].reshape(-1, 2)
        want = np.where(r1[:, 1:2] > r1[:, 0:1], r1 + pair_base, 0)
        assert np.array_equal(ranges[tb:tb + nt], want), k
        assert np.array_equal(rgb[3 * pb:3 * (pb + npix)], s["rgb"]), k
        assert np.array_equal(st["T"][pb:pb + npix], s["T"]), k
        hit = s["T"] < 1.0
        assert np.array_equal(st["last_ids"][pb:pb + npix][hit], s["last_ids"][hit] + pair_base)
        assert abs(losses[k] - s["loss"]) <= 1e-5 * s["loss"], (k, losses[k], s["loss"])
        ok, mabs, mrel = close(gs2_all[:, sb:sb + n].reshape(-1), s["gs2"], atol=1e-7,
                               rtol=1e-4)
        print(f"[parity] multi_raster_grad_{w}x{h}_{n}: max_abs={mabs:.3e} max_rel={mrel:.3e}")
        assert ok, (k, mabs, mrel)
        pair_base += m
    mr.close()


def make_single(w, h, n, seed):
    view, proj, cam = ds.fitter_camera(w, h)
    fit = ds.ImageFitter(w, h, n, view=view, proj=proj, cam_pos=cam, lr=LR, isect_cap=1 << 20)
    fit.set_target(ds.target_image(seed, w, h, 16))
    fit.upload_params(ds.frustum_splats(seed, n, aspect=w / h))
    return fit


def make_multi(sizes, graph):
    inst = []
    for w, h, n, seed in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        inst.append((w, h, n, view, proj, cam))
    mf = ds.MultiFitter(inst, lr=LR, isect_cap=1 << 22)
    for k, (w, h, n, seed) in enumerate(sizes):
        mf.set_target(k, ds.target_image(seed, w, h, 16))
        mf.upload_params(k, ds.frustum_splats(seed, n, aspect=w / h))
    mf.set_graph(graph)
    return mf


@pytest.mark.parametrize("graph", [False, True])
def test_multifit_matches_individual_fits(graph):
    sizes = [s for s in SIZES if s[2] > 0] + [(200, 112, 2500, 25)]
    alone = [make_single(*s) for s in sizes]
    mf = make_multi(sizes, graph)
    # gradients from the same start: equal up to the order of float atomics
    mf.forward_backward()
    for k, a in enumerate(alone):
        a.forward_backward()
        ok, mabs, mrel = close(mf.grads(k), a.grads(), atol=1e-7, rtol=1e-4)
        print(f"[parity] multifit_grads_{sizes[k][0]}x{sizes[k][1]}_{sizes[k][2]}: "
              f"max_abs={mabs:.3e}")
        assert ok, (k, mabs, mrel)
    for f in alone:
        f.step(5)
        f.synchronize()
    mf.step(5)
    mf.synchronize()
    assert mf.step_count() == 5
    mf.forward()
    losses = mf.losses()
    for k, ((w, h, n, seed), a) in enumerate(zip(sizes, alone)):
        ok, mabs, mrel = close(mf.params(k), a.params(), atol=1e-5, rtol=1e-3)
        print(f"[parity] multifit_vs_alone_{w}x{h}_{n} graph={graph}: max_abs={mabs:.3e}")
        assert ok, (k, mabs, mrel)
        a.forward()
        assert abs(a.loss() - losses[k]) <= 1e-4 * a.loss(), (k, a.loss(), losses[k])
        img = mf.image(k)
        assert img.shape == (3 * w * h,) and np.isfinite(img).all()
    mf.close()
    for f in alone:
        f.close()


def test_multifit_graph_replays_match_eager_and_make_progress():
    sizes = [(96, 64, 700, 31), (128, 72, 1200, 32), (48, 48, 200, 33)]
    eager = make_multi(sizes, False)
    graph = make_multi(sizes, True)
    eager.forward()
    l0 = eager.losses()
    tuned = graph.autotune(reps=2)
    assert tuned["fwd_pix_per_thread"] in (1, 2, 4) and tuned["bwd_pix_per_thread"] in (1, 2, 4)
    for _ in range(3):  # replays with host synchronisation in between
        eager.step(10)
        graph.step(10)
        eager.synchronize()
        graph.synchronize()
        for k in range(len(sizes)):
            ok, mabs, _ = close(graph.params(k), eager.params(k), atol=1e-4, rtol=1e-2)
            assert ok, (k, mabs)
    graph.forward()
    l1 = graph.losses()
    assert np.all(np.isfinite(l1)) and np.all(l1 < l0), (l0, l1)
    eager.close()
    graph.close()
