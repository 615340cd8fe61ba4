"""CPU oracle of the splatting hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
module, and only as the checker or the timed CPU baseline: the engine never
calls it. It wraps oracle/liboracle.so (oracle/ds_oracle.cpp, a restatement of
the reference's kernels; see that file for the file:line of each function)
with numpy in/out.

Layouts (field-major, n = splat count):
  params  (14 + 3*bands) * n   pos xyz, log-scale xyz, quat wxyz, sh_dc rgb,
                               sh_rest r,g,b (bands*n each), logit_opacity
  s3      13 * n               pos xyz, cov xx xy xz yy yz zz, rgb, opacity
  s2      10 * n               ndc pos xyz, cov xx xy yy, rgb, opacity
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
        h = C.CDLL(LIB_PATH)
        for suf, T in (("f32", C.c_float), ("f64", C.c_double)):
            getattr(h, f"or_act_fwd_{suf}").argtypes = [_P, C.c_int, C.c_int, _P, _P]
            getattr(h, f"or_act_bwd_{suf}").argtypes = [_P, _P, _P, C.c_int, C.c_int, _P, _P,
                                                        C.c_int]
            getattr(h, f"or_proj_fwd_{suf}").argtypes = [_P, C.c_int, _P, _P]
            getattr(h, f"or_proj_bwd_{suf}").argtypes = [_P, _P, C.c_int, _P, _P, C.c_int]
            getattr(h, f"or_comp_fwd_{suf}").argtypes = [_P, C.c_int, _P, _P, C.c_int, C.c_int,
                                                         _P, _P, _P, _P, _P]
            getattr(h, f"or_comp_bwd_{suf}").argtypes = [_P, C.c_int, _P, _P, C.c_int, C.c_int,
                                                         _P, _P, _P, _P, _P, C.c_int, _P,
                                                         C.c_double, _P]
            getattr(h, f"or_mse_{suf}").argtypes = [_P, _P, C.c_int64, _P]
            getattr(h, f"or_mse_{suf}").restype = T
            getattr(h, f"or_adam_{suf}").argtypes = [_P, _P, _P, _P, C.c_int64, T, T, T, T,
                                                     C.c_int]
        h.or_densify.argtypes = [_P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_float, C.c_float,
                                 C.c_float, C.c_float, C.c_int, C.c_float, C.c_uint64, _P, _P, _P,
                                 _P]
        h.or_densify.restype = C.c_int64
        h.or_tile_assign_sort.argtypes = [_P, C.c_int, C.c_int, C.c_int, C.c_int64, _P, _P,
                                          _P, _P, _P, _P]
        h.or_tile_assign_sort.restype = C.c_int64
        h.or_tile_ranges.argtypes = [_P, C.c_int64, C.c_int, _P]
        h.or_make_key.argtypes = [C.c_uint32, C.c_float]
        h.or_make_key.restype = C.c_uint64
        h.or_set_threads.argtypes = [C.c_int]
        h.or_set_threads.restype = C.c_int
        h.or_fit_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_float,
                                    C.c_float, C.c_float, C.c_int64, _P, _P, C.c_int]
        h.or_fit_create.restype = C.c_void_p
        h.or_fit_destroy.argtypes = [_P]
        h.or_fit_step.argtypes = [_P]
        h.or_fit_step.restype = C.c_float
        h.or_fit_get.argtypes = [_P, _P, _P]
        _lib = h
    return _lib


def _suffix(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def bands(deg: int) -> int:
    return (0, 3, 8, 15)[deg]


def set_threads(n: int) -> int:
    return lib().or_set_threads(n)


def act_fwd(params, n, deg, cam, dtype=np.float64):
    p = _a(params, dtype)
    c = _a(cam, dtype)
    s3 = np.zeros(13 * n, dtype)
    getattr(lib(), f"or_act_fwd_{_suffix(dtype)}")(p.ctypes.data, n, deg, c.ctypes.data,
                                                  s3.ctypes.data)
    return s3


def act_bwd(params, s3, gs3, n, deg, cam, dtype=np.float64, grad_mode=0):
    p, s, g, c = (_a(x, dtype) for x in (params, s3, gs3, cam))
    out = np.zeros((14 + 3 * bands(deg)) * n, dtype)
    getattr(lib(), f"or_act_bwd_{_suffix(dtype)}")(p.ctypes.data, s.ctypes.data, g.ctypes.data,
                                                  n, deg, c.ctypes.data, out.ctypes.data,
                                                  int(grad_mode))
    return out


def proj_fwd(s3, n, pv, dtype=np.float64, init=None):
    s, m = _a(s3, dtype), _a(pv, dtype)
    s2 = np.zeros(10 * n, dtype) if init is None else _a(init, dtype).copy()
    getattr(lib(), f"or_proj_fwd_{_suffix(dtype)}")(s.ctypes.data, n, m.ctypes.data,
                                                   s2.ctypes.data)
    return s2


def proj_bwd(s3, gs2, n, pv, dtype=np.float64, grad_mode=0):
    s, g, m = _a(s3, dtype), _a(gs2, dtype), _a(pv, dtype)
    out = np.zeros(13 * n, dtype)
    getattr(lib(), f"or_proj_bwd_{_suffix(dtype)}")(s.ctypes.data, g.ctypes.data, n,
                                                   m.ctypes.data, out.ctypes.data, int(grad_mode))
    return out


def tile_assign_sort(s2_f32, n, W, H, cap=1 << 25):
    """fp32 transcription: returns dict(keys, ids, n_isects, wanted, visible, rects, counts)."""
    s2 = _a(s2_f32, np.float32)
    ntx, nty = (W + 15) // 16, (H + 15) // 16
    rects = np.zeros(4 * n, np.int32)
    counts = np.zeros(n, np.uint32)
    # first pass sizes the buffers
    wanted = C.c_int64()
    vis = C.c_uint32()
    tmpk = np.zeros(1, np.uint64)
    tmpi = np.zeros(1, np.uint32)
    lib().or_tile_assign_sort(s2.ctypes.data, n, W, H, 0, tmpk.ctypes.data, tmpi.ctypes.data,
                              C.byref(wanted), C.byref(vis), rects.ctypes.data,
                              counts.ctypes.data)
    m = int(min(wanted.value, cap))
    keys = np.zeros(max(m, 1), np.uint64)
    ids = np.zeros(max(m, 1), np.uint32)
    kept = lib().or_tile_assign_sort(s2.ctypes.data, n, W, H, cap, keys.ctypes.data,
                                     ids.ctypes.data, C.byref(wanted), C.byref(vis),
                                     None, None)
    ranges = np.zeros(2 * ntx * nty, np.int32)
    lib().or_tile_ranges(keys.ctypes.data, kept, ntx * nty, ranges.ctypes.data)
    return {"keys": keys[:kept], "ids": ids[:kept], "n_isects": int(kept),
            "wanted": int(wanted.value), "visible": int(vis.value), "rects": rects.reshape(n, 4),
            "counts": counts, "ranges": ranges}


def tile_ranges(keys, ntiles):
    k = _a(keys, np.uint64)
    r = np.zeros(2 * ntiles, np.int32)
    lib().or_tile_ranges(k.ctypes.data, len(k), ntiles, r.ctypes.data)
    return r


def make_key(tile: int, depth: float) -> int:
    return int(lib().or_make_key(tile, depth))


def comp_fwd(s2, n, flat, ranges, W, H, dtype=np.float64):
    s = _a(s2, dtype)
    f = _a(flat, np.uint32)
    r = _a(ranges, np.int32)
    rgb = np.zeros(3 * W * H, dtype)
    T = np.zeros(W * H, dtype)
    last = np.zeros(W * H, np.int32)
    margin = np.zeros(W * H, np.float64)
    e = C.c_int64()
    getattr(lib(), f"or_comp_fwd_{_suffix(dtype)}")(s.ctypes.data, n, f.ctypes.data,
                                                   r.ctypes.data, W, H, rgb.ctypes.data,
                                                   T.ctypes.data, last.ctypes.data,
                                                   margin.ctypes.data, C.byref(e))
    return {"rgb": rgb, "T": T, "last": last, "margin": margin, "E": int(e.value)}


def comp_bwd(s2, n, flat, ranges, W, H, T, last, dLdC, dtype=np.float64, grad_mode=0,
             px_margin=None, amb_margin=5e-5):
    """grad_mode 0: exact dL/dalpha; 1: the reference's formula
    (tile_rasterize_layer.cu:491-495, colour behind not renormalised).

    With px_margin (comp_fwd's "margin"), also returns an allowance per
    gradient element: the summed magnitude of the terms that pixels with a
    decision within amb_margin of its threshold contribute (what a decision
    flip between fp32 and this evaluation may change)."""
    s = _a(s2, dtype)
    f = _a(flat, np.uint32)
    r = _a(ranges, np.int32)
    t = _a(T, dtype)
    l = _a(last, np.int32)
    g = _a(dLdC, dtype)
    gs2 = np.zeros(10 * n, dtype)
    e = C.c_int64()
    pm = _a(px_margin, np.float64) if px_margin is not None else None
    amb = np.zeros(10 * n, dtype) if pm is not None else None
    getattr(lib(), f"or_comp_bwd_{_suffix(dtype)}")(
        s.ctypes.data, n, f.ctypes.data, r.ctypes.data, W, H, t.ctypes.data, l.ctypes.data,
        g.ctypes.data, gs2.ctypes.data, C.byref(e), int(grad_mode),
        pm.ctypes.data if pm is not None else None, C.c_double(amb_margin),
        amb.ctypes.data if amb is not None else None)
    if amb is not None:
        return gs2, int(e.value), amb
    return gs2, int(e.value)


def densify_params(grad_threshold=2e-4, percent_dense=0.01, scene_extent=1.0,
                   min_opacity=0.005, max_scale_frac=0.0, split_scale_div=1.6, seed=0):
    """The engine's decision thresholds on raw parameters (log-scales and
    opacity logits), computed in double and rounded to float as the fitter
    does."""
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    prune_big = max_scale_frac > 0 and scene_extent > 0
    return dict(
        grad_threshold=f32(grad_threshold),
        clone_max_log_scale=f32(math.log(percent_dense * scene_extent))
        if percent_dense > 0 and scene_extent > 0 else -math.inf,
        min_logit=f32(math.log(min_opacity / (1 - min_opacity))) if min_opacity > 0
        else -math.inf,
        prune_max_log_scale=f32(math.log(max_scale_frac * scene_extent)) if prune_big
        else math.inf,
        prune_big=int(prune_big), split_log_div=f32(math.log(split_scale_div)), seed=int(seed))


def densify(slab, m, v, n, deg, accum, visits, p):
    """Clone / split / prune (Kerbl et al. 2023, sec. 5.2) as the engine does
    it, from float parameters, Adam moments and gradient statistics, with the
    thresholds of densify_params. Returns the per-splat actions (0 prune, 1
    keep, 2 clone, 3 split), the new count, and the new slab (float64: split
    samples exact in double), moments."""
    nf = 14 + 3 * bands(deg)
    sl, mm, vv = (_a(x, np.float32) for x in (slab, m, v))
    acc = _a(accum, np.float32)
    vis = _a(visits, np.uint32)
    act = np.zeros(max(n, 1), np.uint8)
    args = [sl.ctypes.data, mm.ctypes.data, vv.ctypes.data, n, nf, acc.ctypes.data,
            vis.ctypes.data, p["grad_threshold"], p["clone_max_log_scale"], p["min_logit"],
            p["prune_max_log_scale"], p["prune_big"], p["split_log_div"], p["seed"],
            act.ctypes.data]
    n_out = lib().or_densify(*args, None, None, None)
    out = np.zeros(nf * n_out, np.float64)
    om = np.zeros(nf * n_out, np.float32)
    ov = np.zeros(nf * n_out, np.float32)
    lib().or_densify(*args, out.ctypes.data, om.ctypes.data, ov.ctypes.data)
    return {"action": act[:n], "n_out": int(n_out), "params": out, "m": om, "v": ov}


def mse(rgb, tgt, npix, dtype=np.float64):
    a, b = _a(rgb, dtype), _a(tgt, dtype)
    g = np.zeros(3 * npix, dtype)
    loss = getattr(lib(), f"or_mse_{_suffix(dtype)}")(a.ctypes.data, b.ctypes.data, npix,
                                                     g.ctypes.data)
    return float(loss), g


def adam(p, g, m, v, lr, b1, b2, eps, step, dtype=np.float64):
    """In place on copies; returns (p, m, v)."""
    p, g, m, v = (_a(x, dtype).copy() for x in (p, g, m, v))
    getattr(lib(), f"or_adam_{_suffix(dtype)}")(p.ctypes.data, g.ctypes.data, m.ctypes.data,
                                               v.ctypes.data, len(p), lr, b1, b2, eps, step)
    return p, m, v


class Fit:
    """fp32 transcription of one full ImageFitter step (the CPU baseline)."""

    def __init__(self, params, target, n, W, H, deg, pv, cam, lr6, b1=0.9, b2=0.999,
                 eps=1e-8, cap=1 << 25, grad_mode=0):
        p, t = _a(params, np.float32), _a(target, np.float32)
        self._keep = (p, t)
        pv, cam, lr6 = _a(pv, np.float32), _a(cam, np.float32), _a(lr6, np.float32)
        self.h = lib().or_fit_create(n, W, H, deg, pv.ctypes.data, cam.ctypes.data,
                                     lr6.ctypes.data, b1, b2, eps, cap, p.ctypes.data,
                                     t.ctypes.data, int(grad_mode))
        self.n, self.W, self.H, self.deg = n, W, H, deg

    def step(self) -> float:
        return float(lib().or_fit_step(self.h))

    def params(self):
        out = np.zeros((14 + 3 * bands(self.deg)) * self.n, np.float32)
        lib().or_fit_get(self.h, out.ctypes.data, None)
        return out

    def image(self):
        out = np.zeros(3 * self.W * self.H, np.float32)
        lib().or_fit_get(self.h, None, out.ctypes.data)
        return out

    def close(self):
        if self.h:
            lib().or_fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
