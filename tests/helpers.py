"""Shared scenes and pipeline compositions for the tests.

The FP64 pipeline below chains the oracle's stages in the reference's step
order (src/pipelines/image_fitter.cpp:103-112): activation -> projection ->
tile assign/sort/ranges -> compositing -> MSE, and back.
"""
from __future__ import annotations

import numpy as np

import oracle

# The reference's finite-difference scene (tests/grad_check.cu:33-37, 72-132):
# 8 splats on a 2x4 grid at staggered depth, anisotropic scales, a 15 degree
# tilt about z, DC colour 0 (-> 0.5), logit 2, a 32x32 target of 0.7, the
# fitter's ortho camera and a camera position of (0, 0, 5).
GC_N, GC_W, GC_H = 8, 32, 32


def ortho(l, r, b, t, n, f):
    m = np.zeros(16, np.float64)
    m[0] = 2 / (r - l)
    m[5] = 2 / (t - b)
    m[10] = -2 / (f - n)
    m[12] = -(r + l) / (r - l)
    m[13] = -(t + b) / (t - b)
    m[14] = -(f + n) / (f - n)
    m[15] = 1
    return m


def gradcheck_scene(deg: int = 0, seed: int = 0):
    n, W, H = GC_N, GC_W, GC_H
    nb = oracle.bands(deg)
    p = np.zeros((14 + 3 * nb) * n, np.float64)
    f = lambda k: p[k * n:(k + 1) * n]  # noqa: E731
    i = np.arange(n)
    f(0)[:] = (i % 4) * 4.0 - 6.0
    f(1)[:] = (i // 4) * 6.0 - 3.0
    f(2)[:] = i * 0.1
    f(3)[:], f(4)[:], f(5)[:] = 1.2, 0.8, 0.5
    f(6)[:], f(7)[:], f(8)[:], f(9)[:] = 0.991, 0.0, 0.0, 0.131
    p[-n:] = 2.0
    if nb:
        rng = np.random.default_rng(seed)
        p[13 * n:13 * n + 3 * nb * n] = rng.normal(0.0, 0.15, 3 * nb * n)
    target = np.full(W * H * 3, 0.7, np.float64)
    hw, hh = W * 0.5, H * 0.5
    zr = min(hw, hh) * 0.25
    pv = ortho(-hw, hw, -hh, hh, -zr, zr)  # identity view
    cam = np.array([0.0, 0.0, 5.0])
    return dict(params=p, n=n, W=W, H=H, deg=deg, target=target, pv=pv, cam=cam)


def random_scene(seed: int = 0, deg: int = 0, n: int = 10, W: int = 40, H: int = 32):
    """Small scene with general rotations and anisotropic scales under a
    perspective camera at (0.2, -0.1, 3) looking at the origin."""
    rng = np.random.default_rng(seed)
    nb = oracle.bands(deg)
    p = np.zeros((14 + 3 * nb) * n, np.float64)
    P = p.reshape(-1, n)
    P[0:2] = rng.uniform(-0.5, 0.5, (2, n))
    P[2] = rng.uniform(-0.4, 0.4, n)
    P[3:6] = rng.normal(-2.2, 0.3, (3, n))
    P[6:10] = rng.normal(0, 1, (4, n))
    P[10:13] = rng.normal(0, 0.6, (3, n))
    if nb:
        p[13 * n:13 * n + 3 * nb * n] = rng.normal(0, 0.2, 3 * nb * n)
    P[-1] = rng.normal(0.5, 1.0, n)
    eye = np.array([0.2, -0.1, 3.0])
    f = -eye / np.linalg.norm(eye)
    s = np.cross(f, [0, 1, 0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = s, u, -f
    V[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    t = np.tan(0.5 * 0.8)
    Pm = np.zeros((4, 4))
    Pm[0, 0], Pm[1, 1] = 1 / (t * W / H), 1 / t
    Pm[2, 2], Pm[2, 3], Pm[3, 2] = -(20 + 0.1) / (20 - 0.1), -2 * 20 * 0.1 / (20 - 0.1), -1
    pv = (Pm @ V).T.reshape(-1)
    target = rng.uniform(0, 1, W * H * 3)
    return dict(params=p, n=n, W=W, H=H, deg=deg, target=target, pv=pv, cam=eye)


def fp64_forward(params, n, deg, cam, pv, W, H, target, pairs=None):
    """FP64 forward. Tile decisions come from the fp32 transcription of the
    (rounded) Splat2D unless `pairs` (a tile_assign_sort result) is given."""
    s3 = oracle.act_fwd(params, n, deg, cam)
    s2 = oracle.proj_fwd(s3, n, pv)
    ta = pairs if pairs is not None else oracle.tile_assign_sort(s2.astype(np.float32), n, W, H)
    comp = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H)
    loss, dLdC = oracle.mse(comp["rgb"], target, W * H)
    return dict(s3=s3, s2=s2, ta=ta, comp=comp, loss=loss, dLdC=dLdC)


def fp64_forward_on(s2_f32, params, n, deg, cam, pv, W, H, target, pairs):
    """FP64 forward whose compositing runs on a given fp32 Splat2D (the GPU's
    own projection output): the discrete decisions of compositing then see
    the same inputs as the GPU, and only the arithmetic differs. The FP64
    activation and projection of `params` are returned too ("s2_fp64"), so
    the fp32 Splat2D can itself be checked against them; the backward of
    the result chains through the FP64 activation state ("s3")."""
    s3 = oracle.act_fwd(params, n, deg, cam)
    s2_64 = oracle.proj_fwd(s3, n, pv)
    s2 = np.asarray(s2_f32, np.float32).astype(np.float64)
    comp = oracle.comp_fwd(s2, n, pairs["ids"], pairs["ranges"], W, H)
    loss, dLdC = oracle.mse(comp["rgb"], target, W * H)
    return dict(s3=s3, s2=s2, s2_fp64=s2_64, ta=pairs, comp=comp, loss=loss, dLdC=dLdC)


def fp64_backward(fw, params, n, deg, cam, pv, W, H, grad_mode=0, amb_margin=None):
    """FP64 backward of fp64_forward. With amb_margin, "gs2_amb" is the
    decision-flip allowance of each dL/dSplat2D element (see oracle.comp_bwd)
    and "amb" flags the splats that have one."""
    r = oracle.comp_bwd(fw["s2"], n, fw["ta"]["ids"], fw["ta"]["ranges"], W, H,
                        fw["comp"]["T"], fw["comp"]["last"], fw["dLdC"], grad_mode=grad_mode,
                        px_margin=fw["comp"]["margin"] if amb_margin is not None else None,
                        amb_margin=amb_margin if amb_margin is not None else 0.0)
    gs2 = r[0]
    gs3 = oracle.proj_bwd(fw["s3"], gs2, n, pv, grad_mode=grad_mode)
    gp = oracle.act_bwd(params, fw["s3"], gs3, n, deg, cam, grad_mode=grad_mode)
    out = dict(gs2=gs2, gs3=gs3, gparams=gp)
    if amb_margin is not None:
        out["gs2_amb"] = r[2]
        out["amb"] = (r[2].reshape(10, n) > 0).any(axis=0)
    return out


def param_allowance(fw, params, n, deg, cam, pv, gs2_amb, grad_mode=0):
    """Propagates a dL/dSplat2D allowance to the parameter gradients. The
    projection and activation backward map each splat's ten Splat2D
    gradients linearly to its parameter gradients; the allowance of a
    parameter gradient is sum_f |d g_param / d g_splat2d_f| * allowance_f,
    with the Jacobian columns found by pushing unit Splat2D gradients
    through the FP64 backward."""
    out = np.zeros(oracle.bands(deg) * 3 * n + 14 * n, np.float64)
    nf = len(out) // n
    amb = np.asarray(gs2_amb, np.float64).reshape(10, n)
    for f in (0, 1, 3, 4, 5, 6, 7, 8, 9):
        if not np.any(amb[f]):
            continue
        e = np.zeros(10 * n, np.float64)
        e[f * n:(f + 1) * n] = 1.0
        gs3 = oracle.proj_bwd(fw["s3"], e, n, pv, grad_mode=grad_mode)
        col = oracle.act_bwd(params, fw["s3"], gs3, n, deg, cam, grad_mode=grad_mode)
        out += np.abs(col) * np.tile(amb[f], nf)
    return out


def fd_grad(scene, h=1e-6, pairs=None):
    """Central finite difference of the FP64 loss w.r.t. every parameter."""
    p0 = scene["params"]
    g = np.zeros_like(p0)
    args = (scene["n"], scene["deg"], scene["cam"], scene["pv"], scene["W"], scene["H"],
            scene["target"])
    for k in range(len(p0)):
        p = p0.copy()
        p[k] += h
        lp = fp64_forward(p, *args, pairs=pairs)["loss"]
        p[k] -= 2 * h
        lm = fp64_forward(p, *args, pairs=pairs)["loss"]
        g[k] = (lp - lm) / (2 * h)
    return g


def field_names(deg):
    nb = oracle.bands(deg)
    names = ["pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z", "rot_w", "rot_x",
             "rot_y", "rot_z", "sh_dc_r", "sh_dc_g", "sh_dc_b"]
    names += [f"sh_rest_{c}[{b}]" for c in "rgb" for b in range(nb)]
    names += ["logit_opacity"]
    return names


def close(a, b, atol=1e-5, rtol=1e-4):
    """|a - b| <= atol + rtol * |b| elementwise; returns (ok, max_abs, max_rel)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    ok = bool(np.all(d <= atol + rtol * np.abs(b)))
    rel = d / np.maximum(np.abs(b), 1e-30)
    return ok, float(d.max(initial=0.0)), float(rel[d > atol].max(initial=0.0))
