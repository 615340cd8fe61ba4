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
    for c, k in [(0, 0), (3, 1), (5, 2), (7, 3), (2, 4), (11, 5), (1, 6), (9, 7), (4, 8), (6, 12)]:
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
        return R.head_fwd(h, g) @ gs

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
