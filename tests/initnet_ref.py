"""numpy reference of the learned initializer (test infrastructure only):
the network of ds_initnet (six 3x3 convolutions: 3 -> 32 /2, 32 -> 32 /2,
32 -> 64 /2, 64 -> 64 /2, 64 -> 64, 64 -> 160; ReLU between; NHWC, GEMM
index k = (kh, kw, ci)), its head (16 splats of 10 values per output cell),
the head's inverse (regression targets of fitted splats) and both backward
passes. With bf16=True it rounds to bf16 (round to nearest even) exactly where
the HIP kernels store bf16: the centred image, the weights, every hidden
activation, and every backward gradient that feeds a GEMM; accumulation stays
in the working precision."""
from __future__ import annotations

import math

import numpy as np

LAYERS = [(3, 32, 3, 2), (32, 32, 3, 2), (32, 64, 3, 2), (64, 64, 3, 2), (64, 64, 3, 1),
          (64, 160, 3, 1)]
SH_C0 = 0.28209479177387814
HEAD = 10  # values per splat
SUB = 4    # splats per output cell side
CLIP = 0.98


def bf16(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(params):
    """Flat params (per layer W[co][kh][kw][ci] then b[co]) -> [(W (co, K), b)]."""
    out, o = [], 0
    for cin, cout, ks, _ in LAYERS:
        k = ks * ks * cin
        w = params[o:o + cout * k].reshape(cout, k)
        o += cout * k
        b = params[o:o + cout]
        o += cout
        out.append((w, b))
    assert o == params.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in layers])


def im2col(x, ks, s):
    """x (H, W, C) -> (Ho*Wo, ks*ks*C) with column index (kh, kw, c)."""
    h, w, c = x.shape
    p = ks // 2
    xp = np.pad(x, ((p, p), (p, p), (0, 0)))
    ho, wo = (h + 2 * p - ks) // s + 1, (w + 2 * p - ks) // s + 1
    cols = np.empty((ho, wo, ks, ks, c), x.dtype)
    for kh in range(ks):
        for kw in range(ks):
            cols[:, :, kh, kw, :] = xp[kh:kh + s * ho:s, kw:kw + s * wo:s, :]
    return cols.reshape(ho * wo, ks * ks * c), ho, wo


def col2im(dcols, shape, ks, s, ho, wo):
    """Adjoint of im2col: (Ho*Wo, ks*ks*C) -> (H, W, C)."""
    h, w, c = shape
    p = ks // 2
    dxp = np.zeros((h + 2 * p, w + 2 * p, c), dcols.dtype)
    d = dcols.reshape(ho, wo, ks, ks, c)
    for kh in range(ks):
        for kw in range(ks):
            dxp[kh:kh + s * ho:s, kw:kw + s * wo:s, :] += d[:, :, kh, kw, :]
    return dxp[p:p + h, p:p + w, :]


def _rounder(bf, dtype):
    if bf:
        return lambda a: bf16(np.asarray(a, np.float32)).astype(dtype)
    return lambda a: np.asarray(a, dtype)


def forward(params, image, W, H, bf=True, dtype=np.float32):
    """image [H*W*3] in [0,1] -> (head values (cells, 160), cache)."""
    rnd = _rounder(bf, dtype)
    layers = unpack(np.asarray(params, dtype))
    x = rnd(np.asarray(image, dtype).reshape(H, W, 3) - 0.5)
    acts = [x]
    a = x
    for i, ((cin, cout, ks, s), (w, b)) in enumerate(zip(LAYERS, layers)):
        cols, ho, wo = im2col(a, ks, s)
        y = cols @ rnd(w).T + b
        if i < len(LAYERS) - 1:
            a = rnd(np.maximum(y, 0)).reshape(ho, wo, cout)
            acts.append(a)
        else:
            h = y
    return h, {"acts": acts, "layers": layers}


def geometry(W, H, fovy=math.radians(60.0)):
    th = math.tan(0.5 * fovy)
    depth = 4.0
    return {"W": W, "H": H, "gw": W // 4, "gh": H // 4, "cell": 4.0, "max_off": 16.0,
            "depth": depth, "max_dz": 1.0, "th": th, "aspect": W / H,
            "ls0": math.log(2.0 * 2.0 * depth * th / H), "max_ls": 2.0, "max_an": 2.0,
            "lsz": -3.0, "logit_bias": 1.0}


ANISO_EPS2 = 1e-8


def splat_order(g):
    """For each splat (anchor row-major), the row of its 9 values in the
    (cells * 16, 9) view of the network's (cells, 144) output."""
    ay, ax = np.divmod(np.arange(g["gw"] * g["gh"]), g["gw"])
    cell = (ay // SUB) * (g["gw"] // SUB) + ax // SUB
    return cell * SUB * SUB + (ay % SUB) * SUB + ax % SUB


def per_splat(h, g):
    """(cells, 160) network output -> (splats, 10) in slab order."""
    return np.asarray(h).reshape(-1, HEAD)[splat_order(g)]


def per_cell(hs, g):
    """Inverse of per_splat."""
    out = np.zeros_like(np.asarray(hs))
    out[splat_order(g)] = hs
    return out.reshape(-1, SUB * SUB * HEAD)


def _anchor_uv(h, g):
    n = g["gw"] * g["gh"]
    c = np.arange(n)
    cx, cy = c % g["gw"], c // g["gw"]
    u = g["cell"] * (cx + 0.5) + g["max_off"] * np.tanh(h[:, 0])
    v = g["cell"] * (cy + 0.5) + g["max_off"] * np.tanh(h[:, 1])
    return 2 * u / g["W"] - 1, 1 - 2 * v / g["H"]


def head_fwd(h, g):
    """head values (splats, 10) in slab order -> SH-0 slab (14 * splats,): the
    splat on the ray through its image position at depth d; mean apparent
    log-scale m, anisotropy vector (h3, h4) = r (cos 2 theta, sin 2 theta)
    with half log-scale difference a = max_an tanh(sqrt(r^2 + eps^2) / max_an)
    and major axis at angle theta; log-scales m +- a shifted by log(d / depth)."""
    n = g["gw"] * g["gh"]
    xn, yn = _anchor_uv(h, g)
    d = g["depth"] + g["max_dz"] * np.tanh(h[:, 9])
    m = g["ls0"] + g["max_ls"] * np.tanh(h[:, 2]) + np.log(d / g["depth"])
    r = np.sqrt(h[:, 3] ** 2 + h[:, 4] ** 2 + ANISO_EPS2)
    a = g["max_an"] * np.tanh(r / g["max_an"])
    phi = 0.25 * np.arctan2(h[:, 4], h[:, 3])
    sig = 1 / (1 + np.exp(-h[:, 5:8]))
    slab = np.concatenate([
        xn * d * g["th"] * g["aspect"], yn * d * g["th"], -d, m + a, m - a,
        m + g["lsz"], np.cos(phi), np.zeros(n), np.zeros(n), np.sin(phi),
        ((sig - 0.5) / SH_C0).T.reshape(-1), h[:, 8] + g["logit_bias"]])
    return slab


def head_bwd(h, g, gs):
    n = g["gw"] * g["gh"]
    G = gs.reshape(14, n)
    d = np.zeros_like(h)
    t0, t1, t2, t9 = (np.tanh(h[:, k]) for k in (0, 1, 2, 9))
    xn, yn = _anchor_uv(h, g)
    dep = g["depth"] + g["max_dz"] * t9
    gx = G[0] * g["th"] * g["aspect"]
    gy = G[1] * g["th"]
    d[:, 0] = gx * dep * (2 / g["W"]) * g["max_off"] * (1 - t0 ** 2)
    d[:, 1] = -gy * dep * (2 / g["H"]) * g["max_off"] * (1 - t1 ** 2)
    gm = G[3] + G[4] + G[5]
    d[:, 2] = gm * g["max_ls"] * (1 - t2 ** 2)
    r2 = h[:, 3] ** 2 + h[:, 4] ** 2 + ANISO_EPS2
    r = np.sqrt(r2)
    ta = np.tanh(r / g["max_an"])
    ga = (G[3] - G[4]) * (1 - ta ** 2) / r
    phi = 0.25 * np.arctan2(h[:, 4], h[:, 3])
    gphi = -np.sin(phi) * G[6] + np.cos(phi) * G[9]
    d[:, 3] = ga * h[:, 3] - 0.25 * gphi * h[:, 4] / r2
    d[:, 4] = ga * h[:, 4] + 0.25 * gphi * h[:, 3] / r2
    d[:, 9] = (gx * xn + gy * yn - G[2] + gm / dep) * g["max_dz"] * (1 - t9 ** 2)
    s = 1 / (1 + np.exp(-h[:, 5:8]))
    d[:, 5:8] = G[10:13].T * s * (1 - s) / SH_C0
    d[:, 8] = G[13]
    return d


def quat_rotation(q):
    """(4, n) quaternions (w, x, y, z), normalised -> (3, 3, n) rotations."""
    w, x, y, z = q / np.linalg.norm(q, axis=0)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def footprint(S, g):
    """(3, n): the fitted splats' covariances projected along the view rays
    and rescaled to depth g["depth"], as (a, b, e) of [[a, b], [b, e]]:
    K Sigma K^T (g["depth"] / d)^2 with K = [[1, 0, x/d], [0, 1, y/d]]."""
    d = np.maximum(-S[2], 1e-3)
    Rm = quat_rotation(S[6:10])
    Sig = np.einsum("ikn,kn,jkn->ijn", Rm, np.exp(2 * S[3:6]), Rm)
    n = S.shape[1]
    K = np.zeros((2, 3, n))
    K[0, 0] = K[1, 1] = 1
    K[0, 2], K[1, 2] = S[0] / d, S[1] / d
    A = np.einsum("ian,abn,jbn->ijn", K, Sig, K) * (g["depth"] / d) ** 2
    return np.stack([A[0, 0], A[0, 1], A[1, 1]])


def head_targets(slab, g):
    """Regression targets (splats, 10) of a fitted slab: image position and
    depth; from the projected footprint's eigenvalues the mean apparent
    log-scale and half their log difference a, from its major axis (angle
    theta) the anisotropy vector's direction (cos 2 theta, sin 2 theta);
    colour; opacity. tanh and sigmoid arguments clamped to CLIP of their
    range."""
    n = g["gw"] * g["gh"]
    S = np.asarray(slab, np.float64).reshape(14, n)
    ay, ax = np.divmod(np.arange(n), g["gw"])
    at = lambda t: np.arctanh(np.clip(t, -CLIP, CLIP))  # noqa: E731
    depth = np.maximum(-S[2], 1e-3)
    xn = S[0] / (depth * g["th"] * g["aspect"])
    yn = S[1] / (depth * g["th"])
    u, v = 0.5 * (xn + 1) * g["W"], 0.5 * (1 - yn) * g["H"]
    t = np.zeros((n, HEAD))
    t[:, 0] = at((u - g["cell"] * (ax + 0.5)) / g["max_off"])
    t[:, 1] = at((v - g["cell"] * (ay + 0.5)) / g["max_off"])
    t[:, 9] = at((depth - g["depth"]) / g["max_dz"])
    a, b, e = footprint(S, g)
    tr, half = 0.5 * (a + e), np.sqrt(0.25 * (a - e) ** 2 + b * b)
    l1 = tr + half
    l2 = np.maximum(tr - half, 1e-12 * l1)
    L1, L2 = 0.5 * np.log(l1), 0.5 * np.log(l2)
    t[:, 2] = at((0.5 * (L1 + L2) - g["ls0"]) / g["max_ls"])
    r = g["max_an"] * at(0.5 * (L1 - L2) / g["max_an"])
    safe = np.where(half > 0, half, 1.0)
    t[:, 3] = r * np.where(half > 0, 0.5 * (a - e) / safe, 1.0)
    t[:, 4] = r * np.where(half > 0, b / safe, 0.0)
    col = np.clip(S[10:13] * SH_C0 + 0.5, 1 - CLIP, CLIP)
    t[:, 5:8] = np.log(col / (1 - col)).T
    t[:, 8] = S[13] - g["logit_bias"]
    return t


def regression(h, t):
    """(loss, dh) of the head regression on (splats, 10) values."""
    d = np.asarray(h, np.float64) - t
    n = d.shape[0]
    return float(np.sum(d * d) / n), 2 * d / n


def backward(cache, h, g, gslab, bf=True, dh=None):
    """Gradients of the flat parameters for the slab gradient gslab (h:
    the network output (cells, 160)), or for a given head-value gradient dh
    (cells, 160)."""
    acts, layers = cache["acts"], cache["layers"]
    dtype = acts[0].dtype
    rnd = _rounder(bf, dtype)
    if dh is None:
        dh = per_cell(head_bwd(per_splat(h, g), g, gslab), g)
    dout = rnd(dh)
    grads = [None] * len(LAYERS)
    for i in range(len(LAYERS) - 1, -1, -1):
        cin, cout, ks, s = LAYERS[i]
        a = acts[i]
        cols, ho, wo = im2col(a, ks, s)
        d2 = dout.reshape(ho * wo, cout)
        grads[i] = (d2.T @ cols, d2.sum(0))
        if i > 0:
            dx = col2im(d2 @ rnd(layers[i][0]), a.shape, ks, s, ho, wo)
            dout = rnd(dx * (a > 0))
    return pack(grads)
