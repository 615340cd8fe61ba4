"""GPU parity of every stage of the hot path against the oracle.

Each stage gets the same inputs on the GPU (through the C-ABI) and in the
oracle. Integer outputs (tile counts, sorted keys, flatten ids, tile ranges)
must be bit-exact with the fp32 transcription; float outputs must satisfy
|gpu - fp64| <= 1e-5 + 1e-4 |fp64| elementwise.

Image pixels where the FP64 oracle took a discrete decision (alpha >= 1/255,
o*g <= 0.99, T <= 1e-4) within a relative margin of 5e-5 can legitimately
differ (fp32 may decide the other way); they are excluded from the image
comparison, counted and bounded. Gradients need no exclusion: a flipped
decision moves a per-splat gradient by about alpha * T * dL/dC, far below the
1e-5 absolute tolerance.
"""
from __future__ import annotations

import json
import os

import ctypes as C

import numpy as np
import pytest

import diffsplat_amd as ds
import oracle
from gpu_stages import Stages, frustum_scene, with_sh
from helpers import close, fp64_backward, fp64_forward, gradcheck_scene

pytestmark = pytest.mark.gpu

ATOL, RTOL = 1e-5, 1e-4
MARGIN = 5e-5
REPORT = os.environ.get("DS_PARITY_REPORT")


def report(name, **kv):
    print(f"[parity] {name}: " + " ".join(f"{k}={v}" for k, v in kv.items()), flush=True)
    if REPORT:
        with open(REPORT, "a") as f:
            f.write(json.dumps({"test": name, **kv}) + "\n")


def check_close(name, gpu, ref, mask=None, atol=ATOL, rtol=RTOL):
    g = np.asarray(gpu, np.float64)
    r = np.asarray(ref, np.float64)
    if mask is not None:
        g, r = g[mask], r[mask]
    ok, mabs, mrel = close(g, r, atol, rtol)
    scale = float(np.abs(r).max(initial=0.0))
    report(name, max_abs=f"{mabs:.3e}", max_rel_above_atol=f"{mrel:.3e}", max_ref=f"{scale:.3e}",
           n=int(g.size))
    assert ok, f"{name}: max_abs={mabs:.3e} max_rel={mrel:.3e}"
    return mabs


SCENES = {
    # name: (seed, n, W, H, ls_mean)
    "c0": (0, 4096, 256, 256, -3.0),
    "ragged": (5, 1500, 200, 120, -2.6),
    "needles": (9, 2000, 200, 120, -2.8),
    # splats covering up to 960 of 1024 tiles: a splat's pair slots span
    # several of the gradient reduction's 256-slot spans
    "wide": (12, 400, 512, 512, -1.2),
}


def scene_params(scene):
    """Splats and camera of a named scene. "needles": half the splats
    stretched 7.4:1 in a random orientation and 40% with opacities between 1
    and 3 times the 1/255 alpha threshold, so that many pixels sit near the
    edge of a splat's alpha-threshold ellipse (the compositing forward skips
    the splats whose box around that ellipse misses a wave's pixels)."""
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = frustum_scene(seed, n, W, H, ls)
    if scene == "needles":
        rng = np.random.default_rng(seed)
        slab = params.reshape(14, n)
        thin = rng.random(n) < 0.5
        slab[3, thin] += 1.0
        slab[4:6, thin] -= 1.0
        faint = rng.random(n) < 0.4
        o = rng.uniform(1.0, 3.0, n) / 255.0
        slab[13, faint] = np.log(o / (1.0 - o))[faint]
        params = slab.reshape(-1).astype(np.float32)
    return params, view, proj, cam


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_activation_projection_forward(deg):
    seed, n, W, H, ls = SCENES["c0"]
    params, view, proj, cam = scene_params("c0")
    cam = np.array([0.3, -0.2, 0.5], np.float32)
    params = with_sh(params, n, deg, seed) if deg else params
    st = Stages(params, n, deg, W, H, view, proj, cam)
    s3 = st.activ_fwd()
    s2 = st.project_fwd()
    s2f = st.activ_project_fwd()
    r3 = oracle.act_fwd(params, n, deg, cam)
    r2 = oracle.proj_fwd(r3, n, st.pv)
    live = np.repeat(r2[2 * n:3 * n] < 1e30, 10)
    check_close(f"activ_fwd_sh{deg}", s3, r3)
    check_close(f"project_fwd_sh{deg}", s2, r2, live)
    check_close(f"activ_project_fwd_sh{deg}", s2f, r2, live)


# c1: the fit workload (one-kernel depth passes); big: above 131072 splats,
# where the depth passes are reduce-then-scan kernels.
# "c1": 30 sort chunks; "chunks64": the most splats sorted by one-kernel passes
# whose last pass writes the pair offsets (64 chunks of 2048); "big": above it
BIG_SCENES = {"c1": (1000, 60000, 1280, 720, -3.0), "chunks64": (1002, 131072, 640, 360, -3.5),
              "big": (1001, 200000, 640, 360, -3.5)}


@pytest.mark.parametrize("scene", ["c0", "ragged", "c1", "chunks64", "big"])
def test_tile_assign_sort_ranges_bitexact(scene):
    seed, n, W, H, ls = BIG_SCENES[scene] if scene in BIG_SCENES else SCENES[scene]
    params, view, proj, cam = frustum_scene(seed, n, W, H, ls)
    st = Stages(params, n, 0, W, H, view, proj, cam)
    s2 = st.activ_project_fwd()
    out = st.raster_fwd()
    ref = oracle.tile_assign_sort(s2, n, W, H)
    report(f"tile_assign_{scene}", n_isects=out["n_isects"], visible=out["visible"])
    assert out["n_isects"] == ref["n_isects"] and out["wanted"] == ref["wanted"]
    assert out["visible"] == ref["visible"]
    assert np.array_equal(out["keys"], ref["keys"])
    assert np.array_equal(out["flatten_ids"], ref["ids"])
    assert np.array_equal(out["tile_ranges"], ref["ranges"])


@pytest.mark.parametrize("W,H,cap", [(1024, 1024, 0), (1024, 1024, 1_001_234), (1040, 1024, 0)])
def test_tile_sort_with_screen_sized_splats(W, H, cap):
    """Splats covering up to every tile: one splat owns whole 4096-slot chunks of
    pairs. 1024x1024 has exactly 4096 tiles (the one-pass tile sort's limit),
    1040x1024 has 4160 (radix passes); the cap ends the pairs mid-chunk."""
    n = 3000
    params, view, proj, cam = frustum_scene(77, n, W, H, -0.8)
    st = Stages(params, n, 0, W, H, view, proj, cam, isect_cap=cap)
    s2 = st.activ_project_fwd()
    out = st.raster_fwd()
    ref = oracle.tile_assign_sort(s2, n, W, H, cap=cap) if cap else oracle.tile_assign_sort(
        s2, n, W, H)
    report(f"tile_sort_big_splats_{W}x{H}_cap{cap}", n_isects=out["n_isects"])
    assert out["n_isects"] == ref["n_isects"] and out["wanted"] == ref["wanted"]
    assert out["visible"] == ref["visible"]
    assert np.array_equal(out["keys"], ref["keys"])
    assert np.array_equal(out["flatten_ids"], ref["ids"])
    assert np.array_equal(out["tile_ranges"], ref["ranges"])


def test_isect_cap_drops_farthest_pairs():
    seed, n, W, H, ls = SCENES["c0"]
    params, view, proj, cam = scene_params("c0")
    cap = 5000
    st = Stages(params, n, 0, W, H, view, proj, cam, isect_cap=cap)
    s2 = st.activ_project_fwd()
    out = st.raster_fwd()
    ref = oracle.tile_assign_sort(s2, n, W, H, cap=cap)
    assert ref["wanted"] > cap and out["n_isects"] == cap == ref["n_isects"]
    assert out["wanted"] == ref["wanted"]
    assert np.array_equal(out["keys"], ref["keys"])
    assert np.array_equal(out["flatten_ids"], ref["ids"])
    assert np.array_equal(out["tile_ranges"], ref["ranges"])


def test_all_culled_renders_black():
    n, W, H = 300, 64, 64
    params, view, proj, cam = frustum_scene(3, n, W, H)
    params = params.copy()
    params[2 * n:3 * n] = 5.0  # behind the camera (camera looks down -z)
    st = Stages(params, n, 0, W, H, view, proj, cam)
    st.activ_project_fwd()
    out = st.raster_fwd()
    assert out["n_isects"] == 0 and out["visible"] == 0
    assert np.all(out["rgb"] == 0.0) and np.all(out["T"] == 1.0)
    gs2 = st.raster_bwd(np.ones(3 * W * H, np.float32))
    assert np.all(gs2 == 0.0)


@pytest.mark.parametrize("n,bits", [(1, 64), (1000, 64), (70001, 44), (300000, 20)])
def test_sort_pairs_u64(n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64) & np.uint64((1 << bits) - 1 if bits < 64
                                                                     else 2 ** 64 - 1)
    keys[: n // 3] = keys[0]  # many ties to check stability
    vals = np.arange(n, dtype=np.uint32)
    dk = ds.DeviceArray.from_numpy(keys)
    dv = ds.DeviceArray.from_numpy(vals)
    ko = ds.DeviceArray(8 * n)
    vo = ds.DeviceArray(4 * n)
    ds.check(ds.lib().ds_sort_pairs_u64(dk.ptr, dv.ptr, ko.ptr, vo.ptr, n, bits, None), "sort")
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(ko.download(np.uint64, n), keys[order])
    assert np.array_equal(vo.download(np.uint32, n), vals[order])


@pytest.mark.parametrize("ppt", [1, 2, 4])
@pytest.mark.parametrize("scene", ["c0", "ragged", "needles"])
def test_composite_forward(scene, ppt):
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = scene_params(scene)
    st = Stages(params, n, 0, W, H, view, proj, cam, pix_per_thread=ppt)
    s2 = st.activ_project_fwd()
    out = st.raster_fwd()
    ref = oracle.comp_fwd(s2, n, out["flatten_ids"], out["tile_ranges"], W, H)
    ok_px = ref["margin"] > MARGIN
    flips = int((~ok_px).sum())
    report(f"composite_fwd_{scene}_ppt{ppt}", ambiguous_px=flips, E_fwd=ref["E"])
    assert flips <= max(3, W * H // 500)
    check_close(f"composite_rgb_{scene}_ppt{ppt}", out["rgb"], ref["rgb"], np.repeat(ok_px, 3))
    check_close(f"composite_T_{scene}_ppt{ppt}", out["T"], ref["T"], ok_px)
    assert np.array_equal(out["last_ids"][ok_px], ref["last"][ok_px])


@pytest.mark.parametrize("ppt", [1, 2, 4])
@pytest.mark.parametrize("grad_mode", [0, 1])
@pytest.mark.parametrize("scene", ["c0", "ragged", "needles", "wide"])
def test_composite_backward(scene, grad_mode, ppt):
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = scene_params(scene)
    target = ds.target_image(seed, W, H, 32)
    st = Stages(params, n, 0, W, H, view, proj, cam, grad_mode=grad_mode, pix_per_thread=ppt)
    s2 = st.activ_project_fwd()
    out = st.raster_fwd(target)
    ref = oracle.comp_fwd(s2, n, out["flatten_ids"], out["tile_ranges"], W, H)
    # fused MSE in the compositing epilogue
    loss_ref, dldc_ref = oracle.mse(ref["rgb"], target, W * H)
    report(f"fused_mse_{scene}", loss_gpu=out["loss"], loss_ref=loss_ref)
    assert abs(out["loss"] - loss_ref) <= 1e-5 + 1e-4 * loss_ref
    ok_px = np.repeat(ref["margin"] > MARGIN, 3)
    check_close(f"fused_mse_grad_{scene}", out["grad_rgb"], dldc_ref, ok_px)
    # backward from the GPU's own dL/dC, oracle from its own FP64 forward state
    gs2 = st.raster_bwd(out["grad_rgb"])
    rg, e_bwd = oracle.comp_bwd(s2, n, out["flatten_ids"], out["tile_ranges"], W, H, ref["T"],
                                ref["last"], out["grad_rgb"], grad_mode=grad_mode)
    report(f"composite_bwd_{scene}_mode{grad_mode}_ppt{ppt}", E_bwd=e_bwd)
    mask = np.ones(10 * n, bool)
    mask[2 * n:3 * n] = False  # pos_z receives no gradient from compositing
    check_close(f"composite_bwd_{scene}_mode{grad_mode}_ppt{ppt}", gs2, rg, mask)


@pytest.mark.parametrize("scene", ["c0", "ragged"])
def test_tile_dispatch_order_leaves_results_unchanged(scene):
    """Heavy-first and row-major compositing dispatch give the same image,
    transmittance, last ids and MSE gradient bit for bit, and the same
    splat gradients up to the order of float atomics; switching the order
    between a forward and its backward is safe."""
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = scene_params(scene)
    target = ds.target_image(seed, W, H, 32)
    outs = []
    for fwd_order, bwd_order in ((1, 1), (0, 0), (0, 1), (1, 0)):
        st = Stages(params, n, 0, W, H, view, proj, cam, pix_per_thread=1)
        ds.check(ds.lib().ds_raster_set_tile_order(st.ras.h, fwd_order), "tile order")
        st.activ_project_fwd()
        out = st.raster_fwd(target)
        ds.check(ds.lib().ds_raster_set_tile_order(st.ras.h, bwd_order), "tile order")
        out["gs2"] = st.raster_bwd(out["grad_rgb"])
        outs.append(out)
    ref = outs[0]
    for out in outs[1:]:
        for k in ("rgb", "T", "last_ids", "grad_rgb", "flatten_ids", "tile_ranges"):
            assert np.array_equal(out[k], ref[k]), k
        ok, mabs, _ = close(out["gs2"], ref["gs2"], atol=1e-7, rtol=1e-5)
        assert ok, mabs


DEPTH_SORT_CASES = {
    # name: (seed, n, W, H, ls_mean): one scene per depth-sort size regime
    "small": (3, 3000, 160, 96, -2.8),
    "many": (4, 200000, 640, 360, -4.0),  # above the one-kernel limit (131072)
}


@pytest.mark.parametrize("case", sorted(DEPTH_SORT_CASES))
def test_depth_sort_modes_give_identical_results(case):
    """Every depth-sort mode the rasterizer takes (one-kernel 8-bit passes up
    to 131072 splats, multi-kernel 8-bit and 11-bit passes) gives the same
    sorted pairs, ranges, image, transmittance and last ids bit for bit, and
    the same gradients up to the order of float sums; the sort-only entry
    leaves the same pairs as a forward."""
    seed, n, W, H, ls = DEPTH_SORT_CASES[case]
    params, view, proj, cam = frustum_scene(seed, n, W, H, ls)
    target = ds.target_image(seed, W, H, 32)
    modes = [m for m in (0, 1, 2) if not (m == 0 and n > 131072)]
    outs = []
    for mode in modes:
        st = Stages(params, n, 0, W, H, view, proj, cam, pix_per_thread=1)
        ds.check(ds.lib().ds_raster_set_depth_sort(st.ras.h, mode), "depth sort")
        st.activ_project_fwd()
        ds.check(ds.lib().ds_raster_sort(st.ras.h, C.byref(st.s2), None), "ds_raster_sort")
        ds.check(ds.lib().ds_synchronize(), "sync")
        sorted_only = st.ras.state()
        out = st.raster_fwd(target)
        assert np.array_equal(sorted_only["flatten_ids"], out["flatten_ids"])
        out["gs2"] = st.raster_bwd(out["grad_rgb"])
        outs.append(out)
    print(f"[parity] depth_sort_{case}: modes {modes}, n_isects {outs[0]['n_isects']}")
    if n > 131072:
        assert ds.lib().ds_raster_set_depth_sort(st.ras.h, 0) != 0  # refused above the limit
    ref = outs[0]
    for out in outs[1:]:
        for k in ("rgb", "T", "last_ids", "grad_rgb", "flatten_ids", "tile_ranges", "keys"):
            if k in ref:
                assert np.array_equal(out[k], ref[k]), k
        ok, mabs, _ = close(out["gs2"], ref["gs2"], atol=1e-7, rtol=1e-5)
        assert ok, mabs


def test_concurrent_sorts_on_several_streams():
    """Eight rasterizers sorting at once on their own streams (every depth-sort
    mode, one-pass and radix tile sorts), three rounds: each round's sorted
    keys, flatten ids and ranges equal the oracle's bit for bit. Kernels of
    several streams sharing the CUs let the waves of one workgroup drift
    apart, which a missing barrier between LDS writes and reads of other
    waves turns into wrong ranks."""
    scenes = [(40 + k, 30000, 640 if k % 2 else 1280, 360 if k % 2 else 1024, -3.6)
              for k in range(8)]
    runs = []
    for k, (seed, n, W, H, ls) in enumerate(scenes):
        params, view, proj, cam = frustum_scene(seed, n, W, H, ls)
        st = Stages(params, n, 0, W, H, view, proj, cam, pix_per_thread=1)
        ds.check(ds.lib().ds_raster_set_depth_sort(st.ras.h, k % 3), "depth sort")
        s2 = st.activ_project_fwd()
        ref = oracle.tile_assign_sort(s2, n, W, H)
        s = C.c_void_p()
        ds.check(ds.lib().ds_stream_create(C.byref(s)), "stream")
        runs.append((st, ref, s))
    try:
        for rnd in range(3):
            for st, _, s in runs:
                ds.check(ds.lib().ds_raster_sort(st.ras.h, C.byref(st.s2), s), "ds_raster_sort")
            for st, ref, s in runs:
                ds.check(ds.lib().ds_stream_synchronize(s), "stream sync")
                out = st.ras.state()
                assert out["n_isects"] == ref["n_isects"], rnd
                assert np.array_equal(out["keys"], ref["keys"]), rnd
                assert np.array_equal(out["flatten_ids"], ref["ids"]), rnd
                assert np.array_equal(out["tile_ranges"], ref["ranges"]), rnd
        print(f"[parity] concurrent_sorts: 8 streams x 3 rounds, pairs "
              f"{[r['n_isects'] for _, r, _ in runs]}")
    finally:
        for _, _, s in runs:
            ds.lib().ds_stream_destroy(s)


@pytest.mark.parametrize("fwd_ppt,bwd_ppt,grad_mode", [(1, 1, 0), (2, 2, 1), (4, 4, 0),
                                                        (1, 2, 0)])
@pytest.mark.parametrize("scene", ["c0", "ragged", "needles"])
def test_fused_forward_backward_matches_separate_calls(scene, fwd_ppt, bwd_ppt, grad_mode):
    """ds_raster_fwd_bwd (one compositing kernel per tile for forward and
    backward when the launch shapes are equal, with the backward skipping the
    pairs the forward found unreached) against ds_raster_fwd + ds_raster_bwd:
    identical image, transmittance, last ids and dL/dC, loss and gradients
    up to the order of float atomics; and against the FP64 oracle."""
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = scene_params(scene)
    target = ds.target_image(seed, W, H, 32)
    st = Stages(params, n, 0, W, H, view, proj, cam, grad_mode=grad_mode)
    ds.check(ds.lib().ds_raster_set_tuning(st.ras.h, fwd_ppt, bwd_ppt), "tuning")
    s2 = st.activ_project_fwd()
    sep = st.raster_fwd(target)
    sep["gs2"] = st.raster_bwd(sep["grad_rgb"])
    d_t = ds.DeviceArray.from_numpy(target)
    d_loss = ds.DeviceArray(4)
    ds.check(ds.lib().ds_splat2d_zero(st.gs2, None), "zero")
    ds.check(ds.lib().ds_raster_fwd_bwd(st.ras.h, st.s2, st.d_rgb.ptr, d_t.ptr, d_loss.ptr,
                                        st.d_grad_rgb.ptr, st.gs2, None), "ds_raster_fwd_bwd")
    fused = st.ras.state()
    fused["rgb"] = st.d_rgb.download(np.float32, 3 * W * H)
    fused["grad_rgb"] = st.d_grad_rgb.download(np.float32, 3 * W * H)
    fused["loss"] = float(d_loss.download(np.float32, 1)[0])
    fused["gs2"] = st.d_gs2.download(np.float32, 10 * n)
    for k in ("rgb", "T", "last_ids", "grad_rgb", "flatten_ids", "tile_ranges"):
        assert np.array_equal(fused[k], sep[k]), k
    assert abs(fused["loss"] - sep["loss"]) <= 1e-6 * sep["loss"]
    ok, mabs, _ = close(fused["gs2"], sep["gs2"], atol=1e-7, rtol=1e-5)
    assert ok, mabs
    ref = oracle.comp_fwd(s2, n, sep["flatten_ids"], sep["tile_ranges"], W, H)
    rg, _ = oracle.comp_bwd(s2, n, sep["flatten_ids"], sep["tile_ranges"], W, H, ref["T"],
                            ref["last"], sep["grad_rgb"], grad_mode=grad_mode)
    mask = np.ones(10 * n, bool)
    mask[2 * n:3 * n] = False
    check_close(f"fused_fwd_bwd_{scene}_ppt{fwd_ppt}{bwd_ppt}_mode{grad_mode}", fused["gs2"], rg,
                mask)


def heavy_first_ok(order, ranges):
    """order is a permutation of the tiles with non-increasing buckets
    floor(log2(pairs)) + 1 (0 for empty tiles)."""
    counts = (ranges[:, 1] - ranges[:, 0]).astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(len(counts)))
    b = np.where(counts > 0, np.floor(np.log2(np.maximum(counts, 1))).astype(np.int64) + 1, 0)
    assert np.all(np.diff(b[order]) <= 0)


@pytest.mark.parametrize("W,H,n", [(256, 256, 4096), (6144, 3456, 6000)])
def test_heavy_first_order_is_a_sorted_permutation(W, H, n):
    """256 tiles keep their buckets in LDS, 82944 tiles re-read them."""
    params, view, proj, cam = frustum_scene(7, n, W, H, -2.0)
    st = Stages(params, n, 0, W, H, view, proj, cam, pix_per_thread=1)
    st.activ_project_fwd()
    out = st.raster_fwd()
    assert "tile_order" in out
    heavy_first_ok(out["tile_order"], out["tile_ranges"].reshape(-1, 2))


@pytest.mark.parametrize("grad_mode", [0, 1])
@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_projection_activation_backward(deg, grad_mode):
    seed, n, W, H, ls = SCENES["c0"]
    params, view, proj, cam = scene_params("c0")
    cam = np.array([0.3, -0.2, 0.5], np.float32)
    params = with_sh(params, n, deg, seed) if deg else params
    st = Stages(params, n, deg, W, H, view, proj, cam, grad_mode=grad_mode)
    s3 = st.activ_fwd()
    st.project_fwd()
    rng = np.random.default_rng(seed + 1)
    gs2 = rng.normal(0, 1e-3, 10 * n).astype(np.float32)
    st.upload_gs2(gs2)
    gs3 = st.project_bwd()
    gp = st.activ_bwd()
    gpf = st.activ_project_bwd()
    r3 = oracle.act_fwd(params, n, deg, cam)
    rgs3 = oracle.proj_bwd(r3, gs2, n, st.pv, grad_mode=grad_mode)
    # the activation backward gates colour on the saved fp32 output; feed both the GPU's
    rgp = oracle.act_bwd(params, s3, gs3, n, deg, cam, grad_mode=grad_mode)
    rgp_chain = oracle.act_bwd(params, r3, rgs3, n, deg, cam, grad_mode=grad_mode)
    tag = f"sh{deg}_mode{grad_mode}"
    check_close(f"project_bwd_{tag}", gs3, rgs3)
    check_close(f"activ_bwd_{tag}", gp, rgp)
    check_close(f"activ_project_bwd_{tag}", gpf, rgp_chain)


def test_mse_and_adam():
    rng = np.random.default_rng(7)
    npix = 123457
    p = rng.uniform(0, 1, 3 * npix).astype(np.float32)
    t = rng.uniform(0, 1, 3 * npix).astype(np.float32)
    dp, dt = ds.DeviceArray.from_numpy(p), ds.DeviceArray.from_numpy(t)
    dl, dg = ds.DeviceArray(4), ds.DeviceArray(4 * 3 * npix)
    ds.check(ds.lib().ds_mse_fwd(dp.ptr, dt.ptr, npix, dl.ptr, None), "mse_fwd")
    ds.check(ds.lib().ds_mse_bwd(dp.ptr, dt.ptr, npix, dg.ptr, None), "mse_bwd")
    loss_ref, g_ref = oracle.mse(p, t, npix)
    loss = float(dl.download(np.float32, 1)[0])
    report("mse", loss=loss, loss_ref=loss_ref)
    assert abs(loss - loss_ref) <= 1e-5 + 1e-4 * loss_ref
    check_close("mse_grad", dg.download(np.float32, 3 * npix), g_ref)
    # Adam: 3 groups with different lr and lr schedules (decay 0 = constant),
    # 5 steps, device step counter
    sizes, lrs, decays = [1000, 77777, 5], [1e-2, 5e-3, 6e-4], [1.0, 0.9, 0.0]
    groups, host = [], []
    for sz, lr, dec in zip(sizes, lrs, decays):
        arr = [rng.normal(0, 1, sz).astype(np.float32) for _ in range(2)]
        dev = [ds.DeviceArray.from_numpy(a) for a in arr] + [ds.DeviceArray(4 * sz)
                                                           for _ in range(2)]
        dev[2].zero()
        dev[3].zero()
        host.append((arr, dev, lr, dec))
        groups.append(ds.AdamGroup(dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, sz, lr, dec))
    tab = (ds.AdamGroup * len(groups))(*groups)
    hp = ds.AdamHParams(0.9, 0.999, 1e-8)
    step = ds.DeviceArray(4)
    step.zero()
    for _ in range(5):
        ds.check(ds.lib().ds_adam_step_dev(tab, len(groups), hp, step.ptr, None), "adam")
    assert int(step.download(np.int32, 1)[0]) == 5
    for (p0, g), dev, lr, dec in host:
        rp, rm, rv = p0.astype(np.float64), np.zeros(len(p0)), np.zeros(len(p0))
        for t_ in range(1, 6):
            lr_t = lr * (dec ** (t_ - 1) if dec > 0 else 1.0)
            rp, rm, rv = oracle.adam(rp, g, rm, rv, lr_t, 0.9, 0.999, 1e-8, t_)
        check_close(f"adam_params_{len(p0)}", dev[0].download(np.float32, len(p0)), rp)
        check_close(f"adam_m_{len(p0)}", dev[2].download(np.float32, len(p0)), rm)
        check_close(f"adam_v_{len(p0)}", dev[3].download(np.float32, len(p0)), rv)


@pytest.mark.parametrize("scene,fused,grad_mode", [("c0", True, 0), ("c0", False, 0),
                                                   ("c0", True, 1), ("c0", False, 1),
                                                   ("wide", True, 0), ("wide", False, 1)])
def test_fitter_step_matches_oracle_chain(scene, fused, grad_mode):
    """One full ImageFitter step: gradients and post-Adam parameters. The step
    stores each splat's summed gradient records into the cleared gradient
    ("wide": sums over runs of pair slots longer than a reduction span)."""
    seed, n, W, H, ls = SCENES[scene]
    params, view, proj, cam = scene_params(scene)
    target = ds.target_image(seed, W, H, 32)
    lr = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3}
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam, lr=lr, fused=fused,
                         grad_mode=grad_mode)
    fit.set_target(target)
    fit.upload_params(params)
    fit.step(1)
    g_gpu = fit.grads()
    p_gpu = fit.params()
    pv = ds.mat_mul(proj, view).astype(np.float64)
    fw = fp64_forward(params.astype(np.float64), n, 0, cam, pv, W, H, target)
    bw = fp64_backward(fw, params.astype(np.float64), n, 0, cam, pv, W, H, grad_mode=grad_mode)
    mask = np.ones(14 * n, bool)
    check_close(f"fitter_grads_{scene}_fused{int(fused)}_mode{grad_mode}", g_gpu, bw["gparams"])
    # post-Adam: first step moves each parameter by about lr * sign(g); compare where
    # |g| is well above the fp32 noise floor of the gradient
    lr_f = np.concatenate([np.full(3 * n, 1e-2), np.full(3 * n, 5e-3), np.full(4 * n, 6e-4),
                           np.full(3 * n, 1e-2), np.full(n, 6e-3)])
    g64 = bw["gparams"]
    m = 0.1 * g64
    v = 0.001 * g64 * g64
    p_ref = params.astype(np.float64) - lr_f * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    strong = mask & (np.abs(g64) > 1e-3 * np.abs(g64).max())
    check_close(f"fitter_adam_{scene}_fused{int(fused)}_mode{grad_mode}", p_gpu, p_ref, strong)


def test_gradcheck_scene_against_finite_differences():
    """The reference's grad_check scene (tests/grad_check.cu:72-132) on the GPU:
    analytic gradients vs FP64 central differences of the oracle's loss."""
    from helpers import fd_grad

    sc = gradcheck_scene(0)
    n, W, H = sc["n"], sc["W"], sc["H"]
    view = np.eye(4, dtype=np.float32).reshape(-1)
    proj = sc["pv"].astype(np.float32)  # identity view: PV = P
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=sc["cam"])
    fit.set_target(sc["target"].astype(np.float32))
    fit.upload_params(sc["params"].astype(np.float32))
    fit.step(1)
    g_gpu = fit.grads()
    fw = fp64_forward(sc["params"], n, 0, sc["cam"], sc["pv"], W, H, sc["target"])
    fd = fd_grad(sc, pairs=fw["ta"])
    check_close("gradcheck_scene_vs_fd", g_gpu, fd, atol=1e-5, rtol=1e-3)
