"""CPU checks of the numpy reference of the learned initializer (the checker
of the GPU initializer tests): its analytic backward against FP64 central
finite differences of its own forward, with the bf16 rounding switched off,
for the head alone and for the whole network."""
from __future__ import annotations

import numpy as np

import initnet_ref as R


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    sizes = sum(cout * (ks * ks * cin + 1) for cin, cout, ks, _ in R.LAYERS)
    return rng.normal(0, scale, sizes)


def test_head_backward_matches_finite_differences():
    g = R.geometry(32, 24)
    n = g["gw"] * g["gh"]
    rng = np.random.default_rng(1)
    h = rng.normal(0, 1, (n, R.HEAD))
    gs = rng.normal(0, 1, 14 * n)
    d = R.head_bwd(h, g, gs)
    eps = 1e-6
    for c, k in [(0, 0), (3, 1), (5, 2), (7, 3), (2, 4), (11, 5), (1, 6), (9, 7), (4, 8), (6, 5),
                 (8, 9), (10, 9)]:
        hp, hm = h.copy(), h.copy()
        hp[c, k] += eps
        hm[c, k] -= eps
        fd = (R.head_fwd(hp, g) @ gs - R.head_fwd(hm, g) @ gs) / (2 * eps)
        assert abs(fd - d[c, k]) <= 1e-7 + 1e-6 * abs(fd), (c, k, fd, d[c, k])


def test_network_backward_matches_finite_differences():
    W = H = 16
    g = R.geometry(W, H)
    rng = np.random.default_rng(2)
    image = rng.uniform(0, 1, W * H * 3)
    p = random_params(3)
    gs = rng.normal(0, 1, 14 * (W // 4) * (H // 4))

    def loss(pp):
        h, _ = R.forward(pp, image, W, H, bf=False, dtype=np.float64)
        return R.head_fwd(R.per_splat(h, g), g) @ gs

    h, cache = R.forward(p, image, W, H, bf=False, dtype=np.float64)
    grad = R.backward(cache, h, g, gs, bf=False)
    assert grad.shape == p.shape
    # a few parameters of every layer (weights and biases)
    offs, o = [], 0
    for cin, cout, ks, _ in R.LAYERS:
        k = ks * ks * cin
        offs += [o + int(rng.integers(0, cout * k)) for _ in range(4)]
        offs += [o + cout * k + int(rng.integers(0, cout))]
        o += cout * (k + 1)
    eps = 1e-6
    for i in offs:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (loss(pp) - loss(pm)) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], np.float32)
    y = R.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0  # tie -> even mantissa
    assert y[2] == np.float32(1.0 + 4 * 2 ** -8)  # tie -> even (rounds up)
    assert y[3] == 1.0 and y[4] == -2.5
    assert abs(y[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_forward_matches_torch_conv2d():
    """The reference's convolution stack (im2col GEMMs over NHWC with the
    (kh, kw, ci) index) against torch.nn.functional.conv2d in float64 on the
    same weights: stride, padding, ReLU placement and weight layout."""
    import torch
    import torch.nn.functional as F

    W, H = 64, 32
    rng = np.random.default_rng(4)
    image = rng.uniform(0, 1, W * H * 3)
    p = random_params(5, scale=0.2)
    h, _ = R.forward(p, image, W, H, bf=False, dtype=np.float64)
    x = torch.tensor(image.reshape(H, W, 3) - 0.5).permute(2, 0, 1)[None]
    for i, ((cin, cout, ks, st), (w, b)) in enumerate(zip(R.LAYERS, R.unpack(p))):
        wt = torch.tensor(w.reshape(cout, ks, ks, cin)).permute(0, 3, 1, 2)
        x = F.conv2d(x, wt, torch.tensor(b), stride=st, padding=ks // 2)
        if i < len(R.LAYERS) - 1:
            x = torch.relu(x)
    want = x[0].permute(1, 2, 0).reshape(-1, R.LAYERS[-1][1]).numpy()
    assert h.shape == want.shape == ((W // 16) * (H // 16), R.SUB * R.SUB * R.HEAD)
    np.testing.assert_allclose(h, want, rtol=1e-10, atol=1e-12)


def fitted_like(rng, g, n):
    """Splats spread in depth, with arbitrary rotations and scales."""
    d = rng.uniform(3.0, 5.0, n)
    xn, yn = rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n)
    S = np.zeros((14, n))
    S[0], S[1], S[2] = xn * d * g["th"] * g["aspect"], yn * d * g["th"], -d
    S[3:6] = g["ls0"] + rng.normal(0, 0.6, (3, n))
    S[6:10] = rng.normal(0, 1, (4, n))
    S[10:14] = rng.normal(0, 1, (4, n))
    return S


def test_footprint_is_the_projected_covariance():
    """footprint() against J Sigma J^T with J the finite-difference Jacobian of
    the map from a 3D point to where its view ray meets the plane at the
    head's depth, and Sigma built from the rotation by rotating a vector."""
    g = R.geometry(64, 48)
    rng = np.random.default_rng(8)
    n = 40
    S = fitted_like(rng, g, n)
    A = R.footprint(S, g)

    def plane(p):
        return g["depth"] * p[:2] / -p[2]

    for c in range(n):
        p = S[0:3, c]
        J = np.zeros((2, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-6
            J[:, k] = (plane(p + e) - plane(p - e)) / 2e-6
        q = S[6:10, c] / np.linalg.norm(S[6:10, c])
        # rotation of the basis vectors by the quaternion, q v q*
        w, u = q[0], q[1:]
        rot = np.stack([2 * np.dot(u, v) * u + (w * w - np.dot(u, u)) * v + 2 * w * np.cross(u, v)
                        for v in np.eye(3)], 1)
        sig = rot @ np.diag(np.exp(2 * S[3:6, c])) @ rot.T
        want = J @ sig @ J.T
        np.testing.assert_allclose([A[0, c], A[1, c], A[2, c]],
                                   [want[0, 0], want[0, 1], want[1, 1]],
                                   rtol=1e-6, atol=1e-9 * np.abs(want).max())


def test_head_targets_invert_the_head():
    """head_targets(head_fwd(h)) == h for head values inside the clamped
    range once the flat splats' small z extent is removed (off the view axis
    the footprint sees it: with it, the footprint values move by a few
    hundredths). Moving a splat along its ray (depth aside) or flipping its
    quaternion's sign changes nothing; swapping its scales (a quarter turn of
    the major axis) negates the anisotropy vector."""
    g = R.geometry(64, 32)
    n = g["gw"] * g["gh"]
    rng = np.random.default_rng(6)
    h = rng.uniform(-1.5, 1.5, (n, R.HEAD))
    h[:, 3:5] = rng.uniform(-2.0, 2.0, (n, 2))
    slab = R.head_fwd(h, g)
    t_flat = R.head_targets(slab, g)
    assert np.abs(t_flat - h)[:, 2:5].max() <= 0.1
    S = slab.reshape(14, n).copy()
    S[5] = -40.0
    t = R.head_targets(S.reshape(-1), g)
    np.testing.assert_allclose(t, h, rtol=0, atol=1e-7)
    f = rng.uniform(0.9, 1.1, n)  # move each splat along its ray
    S[0:3] *= f
    S[3:5] += np.log(f)
    t2 = R.head_targets(S.reshape(-1), g)
    np.testing.assert_allclose(t2[:, :9], h[:, :9], rtol=0, atol=1e-7)
    S[6:10] *= -1  # the quaternion's other sign
    np.testing.assert_allclose(R.head_targets(S.reshape(-1), g), t2, rtol=0, atol=1e-9)
    S[[3, 4]] = S[[4, 3]]  # swap the in-plane scales: the major axis turns
    t3 = R.head_targets(S.reshape(-1), g)
    np.testing.assert_allclose(t3[:, 3:5], -h[:, 3:5], rtol=0, atol=1e-7)
    np.testing.assert_allclose(t3[:, 2], h[:, 2], rtol=0, atol=1e-7)


def test_isotropic_footprint_has_no_anisotropy():
    g = R.geometry(32, 32)
    n = g["gw"] * g["gh"]
    rng = np.random.default_rng(10)
    h = rng.uniform(-1, 1, (n, R.HEAD))
    h[:, 3:5] = 0.0
    S = R.head_fwd(h, g).reshape(14, n)
    S[5] = -40.0
    t = R.head_targets(S.reshape(-1), g)
    assert np.abs(t[:, 3:5]).max() <= 1e-3
    np.testing.assert_allclose(t[:, 2], h[:, 2], atol=1e-7)


def test_shipped_weights_match_the_layout():
    """diffsplat_amd/data/initnet_256.npy: bf16 bit patterns of every weight
    and bias of the six layers, finite; its settings file names the recipe
    and the measurements of its training run."""
    import json
    import os
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "diffsplat_amd", "data")
    bits = np.load(os.path.join(root, "initnet_256.npy"), allow_pickle=False)
    assert bits.dtype == np.uint16
    assert bits.size == sum(cout * (ks * ks * cin + 1) for cin, cout, ks, _ in R.LAYERS)
    w = (bits.astype(np.uint32) << 16).view(np.float32)
    assert np.isfinite(w).all() and np.abs(w).max() > 0
    with open(os.path.join(root, "initnet_256.json")) as f:
        meta = json.load(f)
    assert meta["data"]["pairs"] == 10000 and meta["data"]["fit_steps"] == 1000
    assert meta["params"] == bits.size
    t = meta["time_to_30db"]
    assert t["supervised+render"]["mean_seconds"] < t["plain"]["mean_seconds"]
