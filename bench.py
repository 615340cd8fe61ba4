#!/usr/bin/env python3
"""Fit steps/s of the plain image fitter at 1280x720 with 60,000 splats.

One step = ImageFitter::step of the reference (src/pipelines/image_fitter.cpp:103-112):
zero grads, activation+projection, tile assign + sort + ranges, compositing
with fused MSE, compositing backward, projection/activation backward, Adam.
Inputs are synthetic and resident in HBM before the timed region starts.

Prints ONE JSON line. `roofline` is for the fused compositing forward +
backward (the dominant kernel), timed with HIP events on the fitter's stream
at the state reached after the timed steps; its algorithmic work (splat
evaluations) is counted by the oracle on the same state. `cpu_baseline` is the oracle's fp32
OpenMP transcription of the same step on this host's cores. Beside the line:
`time_to_30db` (the metric's second half: wall time to 30 dB PSNR at
256x256 / 4096 splats from the trained initializer shipped in
diffsplat_amd/data vs from random splats, 512 held-out seeds),
`viewer` (1M-splat 1080p frames/s) and `batched` (64 uneven fits at once).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit steps/s @720p 60k splats; time-to-30dB PSNR learned-init vs plain"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# Algorithmic FP32 work per splat evaluation (FMA = 2 FLOP) of the
# compositing forward (falloff, alpha, colour and transmittance update) and
# backward (falloff, alpha, the colour / alpha / conic gradients).
FWD_FLOP_PER_EVAL = 24.0
BWD_FLOP_PER_EVAL = 60.0
# HBM bytes of one compositing forward + backward (the roofline's kernels:
# k_composite_fused and k_pair_reduce, the API's accumulating form) per
# launch shape "fwd,bwd" (pixels per thread) at two C1 fit states, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes averaged over the 20 roofline
# repetitions at the end of a short eager run, scaled by a 256 MiB copy
# (fetch x2.000, write x1.000): see profiles/fit_720p_60k_counters.md.
# Counters cannot be read live here; the line reports the state whose pair
# count is nearest the measured one.
COMPOSITE_TRAFFIC = {  # state: (pairs, {shape: (read bytes, write bytes)}, how it was reached)
    "early": (1327685, {"1,1": (169.99e6, 72.07e6)},
              "25 steps from random splats (the default bench flags' state), 1.33M pairs"),
    "final": (648137, {"1,1": (120.65e6, 65.00e6)},
              "a 30-step run from the slab of a 2020-step fit, 648k pairs"),
}


def composite_traffic(shape, n_isects):
    """(read, write, state description) of the profiled state nearest in pair
    count, or None when that launch shape was not profiled."""
    name = min(COMPOSITE_TRAFFIC,
               key=lambda k: abs(math.log(max(n_isects, 1) / COMPOSITE_TRAFFIC[k][0])))
    pairs, shapes, how = COMPOSITE_TRAFFIC[name]
    if shape not in shapes:
        return None
    return shapes[shape] + (f"{name} C1 state: {how}",)


CONFIGS = {
    # name: (W, H, N, n_ellipses, seed)
    "c1": (1280, 720, 60000, 128, 1000),
    "c0": (256, 256, 4096, 32, 0),
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS))
    ap.add_argument("--graph", type=int, default=1, help="replay the step as a HIP graph")
    ap.add_argument("--grad-mode", default="exact", choices=["exact", "reference"],
                    help="compositing backward: exact derivative (DS_GRAD_EXACT) or the "
                         "reference kernel's formula (DS_GRAD_REFERENCE)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-single-seconds", type=float, default=10.0,
                    help="seconds of the one-thread CPU figure")
    ap.add_argument("--psnr-every", type=int, default=250,
                    help="timed steps between the PSNR curve's points (one host sync each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-viewer", action="store_true")
    ap.add_argument("--no-time-to-psnr", action="store_true")
    ap.add_argument("--eval-images", type=int, default=512,
                    help="held-out images of the time-to-30dB comparison")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=500,
                    help="steps of every instance in the batched-fitting leg")
    ap.add_argument("--batch-compare", action="store_true",
                    help="batched leg: also the per-instance stream and sequential runs "
                         "(ds_batch, 64 ds_fitters); off by default to keep the run short")
    ap.add_argument("--save-params", default=None, metavar="FILE",
                    help="write the parameter slab after the timed steps (.npy)")
    ap.add_argument("--start-params", default=None, metavar="FILE",
                    help="start the fit from a saved parameter slab (.npy) instead of random "
                         "splats: a later fit state in a short run, for profiling")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (counter calibration)")
    ap.add_argument("--ppt", type=str, default=os.environ.get("DS_PPT", "auto"),
                    help="compositing pixels per thread 'FWD,BWD' (1, 2 or 4) or 'auto'")
    return ap.parse_args()


def roofline_composite(ds, oracle, fit, target, W, H, N, reps=20):
    """The dominant work of the step, the compositing forward + backward,
    timed with HIP events on the fitter's stream over `reps` repetitions on
    the state reached after the timed steps, as the fitter runs it: with
    equal forward and backward launch shapes one k_composite_fused per tile
    (ds_raster_composite_fwd_bwd re-runs it on the last forward's sorted
    pairs), otherwise k_composite_fwd then k_composite_bwd
    (ds_raster_composite_fwd, ds_raster_bwd); both end with k_pair_reduce,
    which adds each splat's per-tile gradient records. Algorithmic work
    counted by the oracle on the same state: 24 FLOP per forward splat
    evaluation, 60 per backward evaluation."""
    L = ds.lib()
    g = ds._lib.Gauss3D()
    s2 = ds._lib.Splat2D()
    gs2 = ds._lib.Splat2D()
    grad_rgb = C.c_void_p()
    ds.check(L.ds_fitter_views(fit.h, C.byref(g), C.byref(s2), C.byref(gs2), C.byref(grad_rgb)),
             "ds_fitter_views")
    ras = C.c_void_p()
    ds.check(L.ds_fitter_get_raster(fit.h, C.byref(ras)), "ds_fitter_get_raster")
    stream = fit.stream()
    fit.forward()
    fit.synchronize()
    fp, bp = C.c_int(), C.c_int()
    ds.check(L.ds_raster_get_tuning(ras, C.byref(fp), C.byref(bp)), "ds_raster_get_tuning")
    fused = fp.value == bp.value
    shape = f"{fp.value},{bp.value}"
    d_target = ds.DeviceArray.from_numpy(np.ascontiguousarray(target, np.float32))
    d_rgb = ds.DeviceArray(4 * 3 * W * H)
    d_loss = ds.DeviceArray(4)
    if fused:
        args = (ras, C.byref(s2), d_rgb.ptr, d_target.ptr, d_loss.ptr, grad_rgb, C.byref(gs2),
                stream)
        run = lambda: ds.check(L.ds_raster_composite_fwd_bwd(*args),  # noqa: E731
                               "ds_raster_composite_fwd_bwd")
        kernels = "k_composite_fused + k_pair_reduce"
    else:
        def run():
            ds.check(L.ds_raster_composite_fwd(ras, C.byref(s2), d_rgb.ptr, stream),
                     "ds_raster_composite_fwd")
            ds.check(L.ds_raster_bwd(ras, C.byref(s2), grad_rgb, C.byref(gs2), stream),
                     "ds_raster_bwd")
        kernels = "k_composite_fwd + k_composite_bwd + k_pair_reduce"
    avg_s = event_ms(ds, stream, run, reps) * 1e-3
    # algorithmic work on the same state, counted by the oracle
    st = fit.raster_state()
    s2_host = ds.download_ptr(s2.pos_x, np.float32, 10 * N)
    g_host = ds.download_ptr(grad_rgb.value, np.float32, 3 * W * H)
    fw = oracle.comp_fwd(s2_host, N, st["flatten_ids"], st["tile_ranges"], W, H,
                         dtype=np.float32)
    _, e_bwd = oracle.comp_bwd(s2_host, N, st["flatten_ids"], st["tile_ranges"], W, H, st["T"],
                               st["last_ids"], g_host, dtype=np.float32)
    e_fwd = int(fw["E"])
    flops = e_fwd * FWD_FLOP_PER_EVAL + e_bwd * BWD_FLOP_PER_EVAL
    achieved = flops / avg_s / 1e12
    traffic = composite_traffic(shape, int(st["n_isects"]))
    return {"kernel": kernels, "launch_shape": shape, "bound": "valu", "unit": "TFLOP/s",
            "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic[0] + traffic[1] if traffic else None,
            "traffic_read_bytes": traffic[0] if traffic else None,
            "traffic_write_bytes": traffic[1] if traffic else None,
            "traffic_state": traffic[2] if traffic else
            f"no counter figures for launch shape {shape}",
            "avg_launch_ms": round(avg_s * 1e3, 4), "splat_evals_fwd": e_fwd,
            "splat_evals_bwd": int(e_bwd), "flop_per_eval_fwd": FWD_FLOP_PER_EVAL,
            "flop_per_eval_bwd": BWD_FLOP_PER_EVAL, "n_isects": int(st["n_isects"]),
            "note": "bound 'valu': the kernel's work is fp32 vector ALU (no matrix-core work), "
                    "priced against the 157.3 TFLOP/s fp32 peak (which counts packed FMA, not "
                    "used here); the kernel is VALU-issue-bound"}


def learned_vs_plain(ds, n_eval=512, target_db=30.0):
    """Time to target_db on held-out synthetic images (256x256, 32 ellipses,
    seeds 90000 + i, none of them a training image), from random splats and
    from the trained initializer shipped with the package (diffsplat_amd/
    data/initnet_256.npy: supervised on 10,000 (image, 1000-step fit) pairs,
    then fine-tuned through the renderer), with the same fitter settings.
    The learned time includes inference and the fallback (the initializer's
    splats are kept only when they render with a lower loss than the random
    ones). The training run's own measurements of the compared recipes
    (supervised only, render loss only) come from its settings file."""
    from diffsplat_amd import initnet_train as T
    net = T.load(T.WEIGHTS)
    with open(T.WEIGHTS.rsplit(".", 1)[0] + ".json") as f:
        meta = json.load(f)
    out = {"plain": T.evaluate(None, n_eval, target_db=target_db),
           "learned": T.evaluate(net, n_eval, target_db=target_db)}
    net.close()
    train = meta["time_to_30db"]
    return {"workload": f"{T.W}x{T.H} synthetic targets, {(T.W // 4) * (T.H // 4)} splats, "
                        f"{n_eval} held-out seeds from {T.EVAL_SEED0}",
            "target_db": target_db, "plain": out["plain"], "learned": out["learned"],
            "speedup": round(out["plain"]["mean_seconds"] / out["learned"]["mean_seconds"], 2),
            "steps_ratio": round(out["plain"]["mean_steps"] / out["learned"]["mean_steps"], 2),
            "recipe": {"pairs": meta["data"]["pairs"], "fit_steps": meta["data"]["fit_steps"],
                       "supervised_epochs": meta["epochs"], "render_steps": meta["render_steps"],
                       "pairs_seconds": meta["data"]["seconds"],
                       "supervised_seconds": meta["train_seconds"],
                       "render_seconds": meta["render_seconds"]},
            "compared_at_training": {
                k: {"mean_seconds": v["mean_seconds"], "mean_steps": v["mean_steps"],
                    "initial_psnr_db": v["initial_psnr_db"], "fits": v["fits"]}
                for k, v in train.items()},
            "lr_pos_decay": T.FIT_LR["lr_pos_decay"],
            "note": "learned time includes the initializer's inference and the random-init "
                    "fallback check; training is a one-off cost (python -m "
                    "diffsplat_amd.initnet_train)"}


def batched_fit(ds, count=64, steps=500, seq_steps=10, compare=True, graph=True):
    """Batched fitting: `count` independent fits of uneven sizes (areas
    log-uniform 128^2..1024x576, splat counts log-uniform 2k..100k, the
    synthetic target and random frustum init of each), `steps` steps each.
    Primary: ds_multifit, every phase of a step one launch over all instances
    and the step one HIP graph. Compared with: the instances' own captured
    steps spread over autotuned streams (ds_batch), and the instances one
    after another on one stream (`seq_steps` batch steps)."""
    lr = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4,
          "lr_opacity": 6e-3}
    sizes = ds.batch_instances(count)
    inst = []
    for w, h, n, seed in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        inst.append((w, h, n, view, proj, cam))
    mf = ds.MultiFitter(inst, lr=lr, isect_cap=1 << 27)
    for k, (w, h, n, seed) in enumerate(sizes):
        mf.set_target(k, ds.target_image(seed, w, h, 32))
        mf.upload_params(k, ds.frustum_splats(seed, n, aspect=w / h))
    mf.forward()
    pairs0 = mf.raster_counts()[0]
    mtune = mf.autotune(reps=3)
    mf.set_graph(graph)
    mf.step(1)
    mf.synchronize()
    t0 = time.perf_counter()
    mf.step(steps)
    mf.synchronize()
    m_el = time.perf_counter() - t0
    mf.forward()
    m_psnr = [ds.psnr_from_loss(float(x)) for x in mf.losses()]
    pairs1 = mf.raster_counts()[0]
    mf.close()
    m_ms = 1e3 * m_el / steps
    out = {"workload": f"{count} uneven fits (128^2..1024x576 pixels, 2k..100k splats), "
                       f"{steps} steps each",
           "instances": count, "steps": steps,
           "pixels": int(sum(w * h for w, h, _, _ in sizes)),
           "splats": int(sum(n for _, _, n, _ in sizes)),
           "instance_steps_per_s": round(count * steps / m_el, 1),
           "ms_per_batch_step": round(m_ms, 4),
           "method": "ds_multifit: one launch per phase over all instances" +
                     (", one HIP graph per step" if graph else ", eager launches"),
           "compositing_tuning": mtune, "pairs_first_step": int(pairs0),
           "pairs_last_step": int(pairs1),
           "psnr_db_mean": round(float(np.mean(m_psnr)), 3),
           "psnr_db_min": round(float(np.min(m_psnr)), 3)}
    if not compare:
        return out

    fits = []
    for w, h, n, seed in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        f = ds.ImageFitter(w, h, n, view=view, proj=proj, cam_pos=cam, lr=lr,
                           isect_cap=min(1 << 25, max(1 << 20, 64 * n)))
        f.set_target(ds.target_image(seed, w, h, 32))
        f.upload_params(ds.frustum_splats(seed, n, aspect=w / h))
        fits.append(f)
    batch = ds.BatchFitter(fits)
    tuned = batch.autotune(reps=3)
    streams = tuned["streams"]
    batch.set_streams(1)
    batch.step(1)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.step(seq_steps)
    batch.synchronize()
    seq_ms = 1e3 * (time.perf_counter() - t0) / seq_steps
    batch.set_streams(streams)
    batch.step(1)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.step(steps)
    batch.synchronize()
    el = time.perf_counter() - t0
    psnr = []
    for f in fits:
        f.forward()
        psnr.append(ds.psnr_from_loss(f.loss()))
    batch.close()
    for f in fits:
        f.close()
    out.update({"streams_batch": {"streams": streams, "seconds": round(el, 4),
                                  "instance_steps_per_s": round(count * steps / el, 1),
                                  "ms_per_batch_step": round(1e3 * el / steps, 4),
                                  "psnr_db_mean": round(float(np.mean(psnr)), 3)},
                "sequential_ms_per_batch_step": round(seq_ms, 4),
                "speedup_vs_sequential": round(seq_ms / m_ms, 2),
                "speedup_vs_streams": round((1e3 * el / steps) / m_ms, 2)})
    return out


def event_ms(ds, stream, fn, reps):
    """Average ms of fn() over reps calls, HIP events on `stream`."""
    L = ds.lib()
    e0, e1 = C.c_void_p(), C.c_void_p()
    ds.check(L.ds_event_create(C.byref(e0)), "event")
    ds.check(L.ds_event_create(C.byref(e1)), "event")
    fn()
    ds.check(L.ds_event_record(e0, stream), "record")
    for _ in range(reps):
        fn()
    ds.check(L.ds_event_record(e1, stream), "record")
    ds.check(L.ds_event_synchronize(e1), "sync")
    ms = C.c_float()
    ds.check(L.ds_event_elapsed_ms(e0, e1, C.byref(ms)), "elapsed")
    L.ds_event_destroy(e0)
    L.ds_event_destroy(e1)
    return ms.value / reps


def viewer_fps(ds, oracle=None, n=1_000_000, W=1920, H=1080, n_cams=120, reps=10):
    """Forward-only renderer on the clustered SH-3 scene over an orbit of
    n_cams cameras at radius 3 (each frame: activation+projection, tile
    assign, sort, compositing). Then, at the last camera: the compositing
    alone (ds_raster_composite_fwd on the frame's sorted pairs) timed with
    HIP events, its share of the frame, and (with the oracle) its algorithmic
    work, 24 FLOP per splat evaluation, against the fp32 peak."""
    params = ds.cluster_scene(2000, n, 3, 64)
    r = ds.SplatRenderer(W, H, n, 3)
    r.upload_params(params)
    cams = ds.orbit_cameras(n_cams, W, H)
    r.render(*cams[0])
    tuning = r.autotune(reps=3)
    r.synchronize()
    t0 = time.perf_counter()
    for view, proj, eye in cams:
        r.render(view, proj, eye)
    r.synchronize()
    el = time.perf_counter() - t0
    st = r.raster_state()
    L = ds.lib()
    stream = r.stream()
    frame_ms = event_ms(ds, stream, lambda: r.render(*cams[-1]), reps)
    s2 = ds._lib.Splat2D()
    ds.check(L.ds_renderer_views(r.h, C.byref(s2)), "ds_renderer_views")
    ras = C.c_void_p()
    ds.check(L.ds_renderer_get_raster(r.h, C.byref(ras)), "ds_renderer_get_raster")
    d_rgb = ds.DeviceArray(4 * 3 * W * H)
    comp_ms = event_ms(ds, stream, lambda: ds.check(
        L.ds_raster_composite_fwd(ras, C.byref(s2), d_rgb.ptr, stream),
        "ds_raster_composite_fwd"), reps)
    out = {"workload": f"viewer {W}x{H}, {n} splats SH degree 3, {n_cams}-camera orbit",
           "frames_per_s": round(n_cams / el, 2), "ms_per_frame": round(1e3 * el / n_cams, 3),
           "n_isects_last_frame": int(st["n_isects"]), "compositing_tuning": tuning,
           "last_camera": {"frame_ms": round(frame_ms, 4), "compositing_ms": round(comp_ms, 4),
                           "compositing_share": round(comp_ms / frame_ms, 4),
                           "project_assign_sort_share": round(1 - comp_ms / frame_ms, 4)}}
    if oracle is not None:
        s2_host = ds.download_ptr(s2.pos_x, np.float32, 10 * n)
        fw = oracle.comp_fwd(s2_host, n, st["flatten_ids"], st["tile_ranges"], W, H,
                             dtype=np.float32)
        e = int(fw["E"])
        tf = e * FWD_FLOP_PER_EVAL / (comp_ms * 1e-3) / 1e12
        out["last_camera"]["compositing_roofline"] = {
            "kernel": "k_composite_fwd", "bound": "valu", "unit": "TFLOP/s",
            "achieved": round(tf, 3), "peak": PEAK_FP32_TFLOPS,
            "frac": round(tf / PEAK_FP32_TFLOPS, 4), "splat_evals": e,
            "flop_per_eval": FWD_FLOP_PER_EVAL}
    r.close()
    return out


def cpu_baseline(oracle, params, target, W, H, N, pv, cam, lr6, seconds, single_seconds):
    """The oracle's fp32 step on every host thread for `seconds`, then on one
    thread for `single_seconds` (a fresh fit from the same start)."""
    def run(threads, secs):
        got = oracle.set_threads(threads)
        fit = oracle.Fit(params, target, N, W, H, 0, pv, cam, lr6)
        for _ in range(2):
            fit.step()
        t0 = time.perf_counter()
        k = 0
        while True:
            fit.step()
            k += 1
            el = time.perf_counter() - t0
            if el >= secs or k >= 200:
                break
        fit.close()
        return got, k, el

    threads, k, el = run(0, seconds)
    _, k1, el1 = run(1, single_seconds)
    oracle.set_threads(0)
    return {"value": round(k / el, 4), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{k} fp32 steps of the same {W}x{H}/{N}-splat fit after 2 warm-up steps "
                      f"(oracle/ds_oracle.cpp, OpenMP, {threads} threads), {el:.1f} s",
            "single_thread": {"value": round(k1 / el1, 4), "unit": "steps/s", "cores": 1,
                              "sample": f"{k1} steps on one thread after 2 warm-up steps, "
                                        f"{el1:.1f} s"}}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on exactly one GPU: every box here has one GPU and this build "
              "measures one GPU's throughput (multi-GPU sharding is design-only).",
              file=sys.stderr)
        sys.exit(2)

    import diffsplat_amd as ds

    if args.copy_probe:
        nbytes = args.copy_probe << 20
        a, b = ds.DeviceArray(nbytes), ds.DeviceArray(nbytes)
        a.zero()
        for _ in range(5):
            ds.check(ds.lib().ds_memcpy_dtod(b.ptr, a.ptr, nbytes, None), "ds_memcpy_dtod")
        ds.check(ds.lib().ds_synchronize(), "ds_synchronize")
        print(json.dumps({"copy_probe_bytes": nbytes, "copies": 5}), flush=True)
        return

    W, H, N, n_ell, seed = CONFIGS[args.config]
    target = ds.target_image(seed, W, H, n_ell)
    params = ds.frustum_splats(seed, N, aspect=W / H)
    if args.start_params:
        params = np.load(args.start_params, allow_pickle=False).astype(np.float32)
        assert params.size == 14 * N, "the saved slab is not this workload's"
    view, proj, cam = ds.fitter_camera(W, H)
    lr = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4,
          "lr_opacity": 6e-3}

    grad_mode = {"exact": 0, "reference": 1}[args.grad_mode]
    fit = ds.ImageFitter(W, H, N, view=view, proj=proj, cam_pos=cam, lr=lr, grad_mode=grad_mode)
    fit.set_target(target)
    fit.upload_params(params)
    if args.ppt == "auto":
        fit.forward()
        tuning = fit.autotune(reps=3)
    else:
        f_ppt, b_ppt = (int(x) for x in args.ppt.split(","))
        fit.set_tuning(f_ppt, b_ppt)
        tuning = {"fwd_pix_per_thread": f_ppt, "bwd_pix_per_thread": b_ppt}
    print(f"[bench] tuning {tuning}", file=sys.stderr, flush=True)
    fit.set_graph(bool(args.graph))
    fit.step(args.warmup)
    # graph mode: the graphs the timed chunks replay are captured now (setup,
    # not steps: nothing runs), so the timed region holds launches only
    chunk = min(args.psnr_every, args.steps) if args.psnr_every > 0 else args.steps
    fit.prepare(chunk)
    if args.steps % chunk:
        fit.prepare(args.steps % chunk)
    fit.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    # the timed steps in chunks of psnr_every: after each chunk the loss of its
    # last step's forward is read (one host sync) for the PSNR curve
    curve = []
    done = 0
    t0 = time.perf_counter()
    while done < args.steps:
        k = min(args.psnr_every, args.steps - done) if args.psnr_every > 0 else args.steps
        fit.step(k)
        done += k
        curve.append([args.warmup + done, round(ds.psnr_from_loss(fit.loss()), 3)])
    fit.synchronize()
    el = time.perf_counter() - t0
    print(f"[bench] {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)
    fit.forward()
    loss = fit.loss()
    value = args.steps / el
    if args.save_params:
        np.save(args.save_params, fit.params(), allow_pickle=False)

    roof = None
    cpu = None
    viewer = None
    ttp = None
    if not args.no_roofline or not args.no_cpu or not args.no_viewer:
        import oracle
    if not args.no_viewer:
        viewer = viewer_fps(ds, oracle if not args.no_roofline else None)
        print(f"[bench] viewer {viewer}", file=sys.stderr, flush=True)
    if not args.no_time_to_psnr:
        ttp = learned_vs_plain(ds, n_eval=args.eval_images)
        print(f"[bench] time to psnr {ttp}", file=sys.stderr, flush=True)
    batched = None
    if not args.no_batch:
        batched = batched_fit(ds, steps=args.batch_steps, compare=args.batch_compare,
                              graph=bool(args.graph))
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)
    if not args.no_roofline:
        roof = roofline_composite(ds, oracle, fit, target, W, H, N)
        print(f"[bench] roofline {roof}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        pv = ds.mat_mul(proj, view)
        lr6 = np.array([lr["lr_pos"], lr["lr_scale"], lr["lr_rot"], lr["lr_sh_dc"], 6e-3,
                        lr["lr_opacity"]], np.float32)
        cpu = cpu_baseline(oracle, params, target, W, H, N, pv, cam, lr6, args.cpu_seconds,
                           args.cpu_single_seconds)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain image fitter {W}x{H}, {N} random-init splats in a 60 deg "
                               f"pinhole frustum, {n_ell}-ellipse synthetic target, "
                               "fwd+bwd+Adam per step",
                   "width": W, "height": H, "splats": N, "graph": bool(args.graph),
                   "compositing_tuning": tuning, "grad_mode": args.grad_mode,
                   "psnr_curve": curve,
                   "loss_after": round(loss, 6),
                   "psnr_db_after": round(ds.psnr_from_loss(loss), 3),
                   "steps_trained": args.steps + args.warmup,
                   "start": "saved slab" if args.start_params else "random splats"},
        "roofline": roof, "cpu_baseline": cpu, "viewer": viewer, "time_to_30db": ttp,
        "batched": batched,
    }
    if cpu:
        out["vs_cpu_baseline"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
