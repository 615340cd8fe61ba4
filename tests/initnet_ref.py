"""numpy reference of the learned initializer (test infrastructure only):
the network of ds_initnet (conv3x3 32 -> conv3x3/2 32 -> conv3x3/2 64 ->
conv3x3 64 -> conv1x1 16, ReLU between, NHWC, GEMM index k = (kh, kw, ci)),
its head and both backward passes. With bf16=True it rounds to bf16 (round
to nearest even) exactly where the HIP kernels store bf16: the centred
image, the weights, every hidden activation, and every backward gradient that
feeds a GEMM; accumulation stays in the working precision."""
from __future__ import annotations

import math

import numpy as np

LAYERS = [(3, 32, 3, 1), (32, 32, 3, 2), (32, 64, 3, 2), (64, 64, 3, 1), (64, 16, 1, 1)]
SH_C0 = 0.28209479177387814
HEAD = 16


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
    """image [H*W*3] in [0,1] -> (head values (cells, 16), cache)."""
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
    return {"W": W, "H": H, "gw": W // 4, "gh": H // 4, "cell": 4.0, "max_off": 4.0,
            "depth": depth, "th": th, "aspect": W / H,
            "ls0": math.log(2.0 * 2.0 * depth * th / H), "max_ls": 1.5, "lsz": -1.0,
            "logit_bias": 1.0}


def head_fwd(h, g):
    """head values (cells, 16) -> SH-0 slab (14 * cells,)."""
    n = g["gw"] * g["gh"]
    c = np.arange(n)
    cx, cy = c % g["gw"], c // g["gw"]
    u = g["cell"] * (cx + 0.5) + g["max_off"] * np.tanh(h[:, 0])
    v = g["cell"] * (cy + 0.5) + g["max_off"] * np.tanh(h[:, 1])
    xn, yn = 2 * u / g["W"] - 1, 1 - 2 * v / g["H"]
    sx = g["ls0"] + g["max_ls"] * np.tanh(h[:, 2])
    sy = g["ls0"] + g["max_ls"] * np.tanh(h[:, 3])
    th = 0.5 * h[:, 4]
    sig = 1 / (1 + np.exp(-h[:, 5:8]))
    slab = np.concatenate([
        xn * g["depth"] * g["th"] * g["aspect"], yn * g["depth"] * g["th"],
        np.full(n, -g["depth"]), sx, sy, 0.5 * (sx + sy) + g["lsz"],
        np.cos(th), np.zeros(n), np.zeros(n), np.sin(th),
        ((sig - 0.5) / SH_C0).T.reshape(-1), h[:, 8] + g["logit_bias"]])
    return slab


def head_bwd(h, g, gs):
    n = g["gw"] * g["gh"]
    G = gs.reshape(14, n)
    d = np.zeros_like(h)
    t0, t1, t2, t3 = (np.tanh(h[:, k]) for k in range(4))
    d[:, 0] = G[0] * g["depth"] * g["th"] * g["aspect"] * (2 / g["W"]) * g["max_off"] * (1 - t0 ** 2)
    d[:, 1] = -G[1] * g["depth"] * g["th"] * (2 / g["H"]) * g["max_off"] * (1 - t1 ** 2)
    gz = 0.5 * G[5]
    d[:, 2] = (G[3] + gz) * g["max_ls"] * (1 - t2 ** 2)
    d[:, 3] = (G[4] + gz) * g["max_ls"] * (1 - t3 ** 2)
    th = 0.5 * h[:, 4]
    d[:, 4] = 0.5 * (-np.sin(th) * G[6] + np.cos(th) * G[9])
    s = 1 / (1 + np.exp(-h[:, 5:8]))
    d[:, 5:8] = G[10:13].T * s * (1 - s) / SH_C0
    d[:, 8] = G[13]
    return d


def backward(cache, h, g, gslab, bf=True):
    """Gradients of the flat parameters for the slab gradient gslab."""
    acts, layers = cache["acts"], cache["layers"]
    dtype = acts[0].dtype
    rnd = _rounder(bf, dtype)
    dout = rnd(head_bwd(h, g, gslab))
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
