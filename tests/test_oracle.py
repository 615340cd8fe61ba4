"""Pins the CPU oracle (oracle/ds_oracle.cpp) against checks that do not
restate its formulas: scipy's rotations and spherical harmonics, finite
differences, brute force on small inputs, closed forms and invariants, and the
reference's own finite-difference scene (tests/grad_check.cu:72-132)."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation
from scipy.special import sph_harm_y

import oracle
from helpers import (GC_N, fd_grad, field_names, fp64_backward, fp64_forward, random_scene,
                     gradcheck_scene, ortho)

rng0 = np.random.default_rng(1234)


def slab(n, deg=0, rng=rng0):
    nb = oracle.bands(deg)
    p = np.zeros((14 + 3 * nb) * n)
    f = lambda k: p[k * n:(k + 1) * n]  # noqa: E731
    for k in range(3):
        f(k)[:] = rng.uniform(-1, 1, n)
    for k in range(3, 6):
        f(k)[:] = rng.normal(-1.5, 0.4, n)
    q = rng.normal(0, 1, (4, n))
    for k in range(4):
        f(6 + k)[:] = q[k]
    for k in range(10, 13):
        f(k)[:] = rng.normal(0, 0.5, n)
    if nb:
        p[13 * n:13 * n + 3 * nb * n] = rng.normal(0, 0.2, 3 * nb * n)
    p[-n:] = rng.normal(0, 1, n)
    return p


def cov_of(s3, n):
    c = s3.reshape(13, n)
    return np.stack([np.array([[c[3, i], c[4, i], c[5, i]], [c[4, i], c[6, i], c[7, i]],
                               [c[5, i], c[7, i], c[8, i]]]) for i in range(n)])


# ------------------------------------------------------------------ activation
def test_covariance_matches_scipy_rotation():
    n = 64
    p = slab(n)
    s3 = oracle.act_fwd(p, n, 0, np.zeros(3))
    P = p.reshape(-1, n)
    q = P[6:10].T  # w, x, y, z
    R = Rotation.from_quat(np.c_[q[:, 1:], q[:, 0]]).as_matrix()  # scipy wants x,y,z,w
    S = np.exp(P[3:6].T)
    ref = np.einsum("nij,nj,nkj->nik", R, S * S, R)
    np.testing.assert_allclose(cov_of(s3, n), ref, rtol=1e-12, atol=1e-14)


def test_identity_quaternion_gives_diagonal_covariance():
    n = 5
    p = slab(n)
    P = p.reshape(-1, n)
    P[6], P[7:10] = 3.0, 0.0  # unnormalised identity
    s3 = oracle.act_fwd(p, n, 0, np.zeros(3)).reshape(13, n)
    np.testing.assert_allclose(s3[[3, 6, 8]], np.exp(2 * P[3:6]), rtol=1e-14)
    np.testing.assert_array_equal(s3[[4, 5, 7]], 0.0)


def test_zero_quaternion_is_guarded():
    n = 1
    p = slab(n)
    p.reshape(-1, n)[6:10] = 0.0
    s3 = oracle.act_fwd(p, n, 0, np.zeros(3))
    assert np.all(np.isfinite(s3))


def test_opacity_sigmoid_and_dc_colour():
    n = 7
    p = slab(n)
    s3 = oracle.act_fwd(p, n, 0, np.zeros(3)).reshape(13, n)
    P = p.reshape(-1, n)
    np.testing.assert_allclose(s3[12], 1 / (1 + np.exp(-P[13])), rtol=1e-14)
    col = np.clip(0.28209479177387814 * P[10:13] + 0.5, 0, 1)
    np.testing.assert_allclose(s3[9:12], col, rtol=1e-14)


def real_sh_3dgs(l, m, v):
    """3DGS basis from scipy's complex SH (which carries the Condon-Shortley
    phase): sqrt2 Im Y_l^|m| (m < 0), Y_l^0, sqrt2 Re Y_l^m (m > 0)."""
    theta = np.arccos(np.clip(v[2], -1, 1))
    phi = np.arctan2(v[1], v[0])
    Y = sph_harm_y(l, abs(m), theta, phi)
    if m < 0:
        return math.sqrt(2) * Y.imag
    if m == 0:
        return Y.real
    return math.sqrt(2) * Y.real


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_sh_basis_matches_scipy(deg):
    """Each band alone (coefficient 0.1, DC 0) at random view directions; the
    colour minus 0.5 divided by 0.1 is that basis function."""
    nb = oracle.bands(deg)
    rng = np.random.default_rng(deg)
    for _ in range(8):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        n = nb
        p = np.zeros((14 + 3 * nb) * n)
        P = p.reshape(-1, n)
        P[6] = 1.0  # identity rotation
        for k in range(nb):
            p[13 * n + k * n + k] = 0.1  # red channel, band k, splat k
        s3 = oracle.act_fwd(p, n, deg, v).reshape(13, n)  # splats at the origin
        got = (s3[9] - 0.5) / 0.1
        lm = [(l, m) for l in range(1, deg + 1) for m in range(-l, l + 1)]
        want = np.array([real_sh_3dgs(l, m, v) for l, m in lm])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ projection
def lookat(eye, ctr, up):
    f = np.asarray(ctr, float) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    M = np.eye(4)
    M[0, :3], M[1, :3], M[2, :3] = s, u, -f
    M[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    return M


def perspective(fovy, aspect, n, f):
    t = math.tan(fovy / 2)
    M = np.zeros((4, 4))
    M[0, 0], M[1, 1] = 1 / (aspect * t), 1 / t
    M[2, 2], M[2, 3], M[3, 2] = -(f + n) / (f - n), -2 * f * n / (f - n), -1
    return M


def ndc_of(M, x):
    c = M @ np.append(x, 1.0)
    return c[:3] / c[3]


def test_projection_jacobian_and_covariance_by_finite_differences():
    n = 32
    rng = np.random.default_rng(3)
    M = perspective(0.9, 1.6, 0.1, 50) @ lookat(np.array([0.5, 0.3, 4.0]), [0, 0, 0], [0, 1, 0])
    pv = M.T.reshape(-1)  # column-major
    s3 = oracle.act_fwd(slab(n, rng=rng), n, 0, np.zeros(3))
    s2 = oracle.proj_fwd(s3, n, pv).reshape(10, n)
    C3 = cov_of(s3, n)
    X = s3.reshape(13, n)[:3].T
    for i in range(n):
        np.testing.assert_allclose(s2[:3, i], ndc_of(M, X[i]), rtol=1e-12)
        J = np.zeros((2, 3))
        for k in range(3):
            h = np.zeros(3)
            h[k] = 1e-6
            J[:, k] = (ndc_of(M, X[i] + h)[:2] - ndc_of(M, X[i] - h)[:2]) / 2e-6
        S2 = J @ C3[i] @ J.T
        np.testing.assert_allclose(s2[[3, 4, 5], i], [S2[0, 0], S2[0, 1], S2[1, 1]],
                                   rtol=1e-6, atol=1e-12)
        np.testing.assert_array_equal(s2[6:, i], s3.reshape(13, n)[9:, i])


def test_projection_culls_behind_camera_and_zeroes_gradients():
    n = 3
    p = slab(n)
    s3 = oracle.act_fwd(p, n, 0, np.zeros(3)).reshape(13, n)
    s3[2] = [-3.0, 2.0, -1.0]  # the middle one is behind a camera looking down -z
    M = perspective(1.0, 1.0, 0.1, 10)
    pv = M.T.reshape(-1)
    s2 = oracle.proj_fwd(s3.reshape(-1), n, pv).reshape(10, n)
    assert s2[2, 1] == np.float32(3.4028234663852886e38)
    g = oracle.proj_bwd(s3.reshape(-1), np.ones(10 * n), n, pv).reshape(13, n)
    assert np.all(g[:, 1] == 0) and np.any(g[:, 0] != 0)


def test_orthographic_projection_is_linear():
    n = 6
    M = ortho(-16, 16, -9, 9, -4, 4).reshape(4, 4).T
    s3 = oracle.act_fwd(slab(n), n, 0, np.zeros(3))
    s2 = oracle.proj_fwd(s3, n, M.T.reshape(-1)).reshape(10, n)
    J = M[:2, :3]
    for i, C in enumerate(cov_of(s3, n)):
        S2 = J @ C @ J.T
        np.testing.assert_allclose(s2[[3, 4, 5], i], [S2[0, 0], S2[0, 1], S2[1, 1]], rtol=1e-12)


# ------------------------------------------------------------------ tile assignment
def s2_from(x, y, z, cxx, cxy, cyy, n):
    s2 = np.zeros((10, n), np.float32)
    s2[0], s2[1], s2[2], s2[3], s2[4], s2[5] = x, y, z, cxx, cxy, cyy
    s2[6:9] = 0.5
    s2[9] = 0.8
    return s2.reshape(-1)


def test_single_splat_rectangle_closed_form():
    W, H = 160, 96  # 10 x 6 tiles
    # centre at pixel (80, 48), sigma_x = 5 px, sigma_y = 3 px -> 3 sigma = 15 / 9 px
    sx, sy = 5 * 2 / W, 3 * 2 / H
    s2 = s2_from(0.0, 0.0, 0.5, sx * sx, 0.0, sy * sy, 1)
    r = oracle.tile_assign_sort(s2, 1, W, H)
    # x: pixels 65..95 -> tiles 4..5; y: pixels 39..57 -> tiles 2..3
    assert r["rects"][0].tolist() == [4, 5, 2, 3]
    assert r["n_isects"] == 4
    assert [k >> 32 for k in r["keys"].tolist()] == [2 * 10 + 4, 2 * 10 + 5, 3 * 10 + 4,
                                                     3 * 10 + 5]


@pytest.mark.parametrize("case", ["subpixel", "offscreen", "det_le_0", "far", "behind"])
def test_culling_cases(case):
    W, H = 64, 64
    args = dict(x=0.0, y=0.0, z=0.2, cxx=1e-3, cxy=0.0, cyy=1e-3)
    if case == "subpixel":
        args.update(cxx=1e-7, cyy=1e-7)
    elif case == "offscreen":
        args.update(x=1.5)
    elif case == "det_le_0":
        args.update(cxy=2e-3)
    elif case == "far":
        args.update(z=1.01)
    elif case == "behind":
        args.update(z=3.4028234663852886e38)
    r = oracle.tile_assign_sort(s2_from(n=1, **args), 1, W, H)
    assert r["visible"] == 0 and r["n_isects"] == 0 and np.all(r["ranges"] == 0)


def test_rectangles_cover_exactly_the_3sigma_box_brute_force():
    """Tiles whose pixel square meets the 3-sigma box, in FP64, match the
    fp32 rectangle except where a box edge lies within 1e-4 px of a tile edge."""
    rng = np.random.default_rng(11)
    n, W, H = 400, 200, 136
    x, y = rng.uniform(-1.1, 1.1, n), rng.uniform(-1.1, 1.1, n)
    cxx, cyy = rng.uniform(1e-5, 3e-3, n), rng.uniform(1e-5, 3e-3, n)
    cxy = rng.uniform(-0.9, 0.9, n) * np.sqrt(cxx * cyy)
    s2 = s2_from(x, y, rng.uniform(-0.9, 0.9, n), cxx, cxy, cyy, n)
    r = oracle.tile_assign_sort(s2, n, W, H)
    S = s2.reshape(10, n).astype(np.float64)
    ntx, nty = (W + 15) // 16, (H + 15) // 16
    checked = 0
    for i in range(n):
        ex, ey = 3 * math.sqrt(S[3, i]), 3 * math.sqrt(S[5, i])
        px0, px1 = (S[0, i] - ex + 1) / 2 * W, (S[0, i] + ex + 1) / 2 * W
        py0, py1 = (1 - S[1, i] - ey) / 2 * H, (1 - S[1, i] + ey) / 2 * H
        edges = np.array([px0, px1, py0, py1]) / 16
        if np.any(np.abs(edges - np.round(edges)) < 1e-4 / 16):
            continue
        tx0, tx1 = max(math.floor(px0 / 16), 0), min(math.floor(px1 / 16), ntx - 1)
        ty0, ty1 = max(math.floor(py0 / 16), 0), min(math.floor(py1 / 16), nty - 1)
        big = max(ex, ey) * 2 >= max(2 / W, 2 / H)
        vis = tx0 <= tx1 and ty0 <= ty1 and abs(S[0, i]) <= 1 and abs(S[1, i]) <= 1 and big
        want = [tx0, tx1, ty0, ty1] if vis else [0, -1, 0, -1]
        assert r["rects"][i].tolist() == want, i
        checked += 1
    assert checked > 300


def test_keys_sort_and_ranges_brute_force():
    rng = np.random.default_rng(5)
    n, W, H = 300, 128, 80
    z = rng.uniform(-1, 1, n).astype(np.float32)
    z[:40] = z[0]  # equal depths: ties must keep splat order
    s2 = s2_from(rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), z,
                 rng.uniform(1e-4, 2e-3, n), 0.0, rng.uniform(1e-4, 2e-3, n), n)
    r = oracle.tile_assign_sort(s2, n, W, H)
    ntx = (W + 15) // 16
    pairs = []
    for i in range(n):
        tx0, tx1, ty0, ty1 = r["rects"][i]
        for ty in range(ty0, ty1 + 1):
            for tx in range(tx0, tx1 + 1):
                pairs.append((ty * ntx + tx, float(z[i]), i))
    pairs.sort(key=lambda t: (t[0], t[1], t[2]))
    assert r["n_isects"] == len(pairs)
    assert r["ids"].tolist() == [p[2] for p in pairs]
    assert [int(k) >> 32 for k in r["keys"]] == [p[0] for p in pairs]
    ntiles = ntx * ((H + 15) // 16)
    for t in range(ntiles):
        idx = [k for k, p in enumerate(pairs) if p[0] == t]
        want = [idx[0], idx[-1] + 1] if idx else [0, 0]
        assert r["ranges"][2 * t:2 * t + 2].tolist() == want


def test_depth_key_orders_like_float():
    z = np.concatenate([np.linspace(-1, 1, 101), [0.0, -0.0, 1e-30, -1e-30]]).astype(np.float32)
    keys = np.array([oracle.make_key(7, float(v)) for v in z], np.uint64)
    assert np.all(keys >> np.uint64(32) == 7)
    low = keys & np.uint64(0xFFFFFFFF)
    order = np.argsort(low, kind="stable")
    assert np.all(np.diff(z[order]) >= 0)
    assert oracle.make_key(0, 0.5) & 0xFFFFFFFF == (0x3F000000 ^ 0x80000000)
    assert oracle.make_key(0, -0.5) & 0xFFFFFFFF == (~0xBF000000) & 0xFFFFFFFF


def test_isect_cap_keeps_the_nearest_pairs():
    rng = np.random.default_rng(9)
    n, W, H = 200, 96, 96
    s2 = s2_from(rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                 rng.uniform(1e-3, 5e-3, n), 0.0, rng.uniform(1e-3, 5e-3, n), n)
    full = oracle.tile_assign_sort(s2, n, W, H)
    cap = full["n_isects"] // 2
    part = oracle.tile_assign_sort(s2, n, W, H, cap=cap)
    assert part["n_isects"] == cap and part["wanted"] == full["n_isects"]
    zs = s2.reshape(10, n)[2]
    kept_depth = zs[part["ids"]].max()
    dropped = set(full["ids"].tolist()) - set(part["ids"].tolist())
    assert all(zs[d] >= kept_depth for d in dropped)


# ------------------------------------------------------------------ compositing
def composite_brute(s2, n, flat, ranges, W, H):
    """Per-pixel loop in pure Python (small inputs only)."""
    S = np.asarray(s2, np.float64).reshape(10, n)
    ntx = (W + 15) // 16
    rgb = np.zeros((H, W, 3))
    Tm = np.ones((H, W))
    for i in range(H):
        for j in range(W):
            t = (i // 16) * ntx + j // 16
            px = np.float32(2) * (np.float32(j) + np.float32(0.5)) / np.float32(W) - np.float32(1)
            py = np.float32(1) - np.float32(2) * (np.float32(i) + np.float32(0.5)) / np.float32(H)
            T = 1.0
            for k in range(ranges[2 * t], ranges[2 * t + 1]):
                g = flat[k]
                Cm = np.array([[S[3, g], S[4, g]], [S[4, g], S[5, g]]])
                d = np.array([S[0, g] - float(px), S[1, g] - float(py)])
                a = min(0.99, S[9, g] * math.exp(-0.5 * d @ np.linalg.solve(Cm, d)))
                if a < 1 / 255:
                    continue
                rgb[i, j] += S[6:9, g] * a * T
                T *= 1 - a
                if T <= 1e-4:
                    break
            Tm[i, j] = T
    return rgb.reshape(-1), Tm.reshape(-1)


def small_scene(n=12, W=40, H=24, seed=2):
    rng = np.random.default_rng(seed)
    s2 = s2_from(rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n),
                 rng.uniform(-0.9, 0.9, n), rng.uniform(2e-3, 2e-2, n),
                 rng.uniform(-1e-3, 1e-3, n), rng.uniform(2e-3, 2e-2, n), n)
    s = s2.reshape(10, n)
    s[6:9] = rng.uniform(0, 1, (3, n))
    s[9] = rng.uniform(0.3, 0.98, n)
    return s.reshape(-1), n, W, H


def test_compositing_matches_brute_force():
    s2, n, W, H = small_scene()
    ta = oracle.tile_assign_sort(s2, n, W, H)
    out = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H)
    rgb, T = composite_brute(s2, n, ta["ids"], ta["ranges"], W, H)
    np.testing.assert_allclose(out["rgb"], rgb, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(out["T"], T, rtol=1e-10, atol=1e-12)


def test_compositing_weights_sum_to_one_minus_T():
    s2, n, W, H = small_scene(n=30, seed=4)
    s = s2.reshape(10, n)
    s[6:9] = 1.0  # white splats: C = sum alpha_i T_i = 1 - T_final
    ta = oracle.tile_assign_sort(s2, n, W, H)
    out = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H)
    np.testing.assert_allclose(out["rgb"].reshape(-1, 3), np.repeat(1 - out["T"], 3).reshape(
        -1, 3), rtol=0, atol=1e-14)


def test_single_isotropic_splat_closed_form():
    W = H = 32
    sig = 4 * 2 / W  # 4 px in NDC
    s2 = s2_from(0.0, 0.0, 0.1, sig * sig, 0.0, sig * sig, 1).reshape(10, 1)
    s2[6:9, 0] = [0.2, 0.5, 0.9]
    s2[9, 0] = 0.7
    ta = oracle.tile_assign_sort(s2.reshape(-1), 1, W, H)
    out = oracle.comp_fwd(s2.reshape(-1), 1, ta["ids"], ta["ranges"], W, H)
    rgb = out["rgb"].reshape(H, W, 3)
    for (i, j) in [(15, 15), (16, 20), (10, 14), (2, 2)]:
        dx, dy = (j + 0.5 - 16) / 4, (i + 0.5 - 16) / 4
        a = min(0.99, 0.7 * math.exp(-0.5 * (dx * dx + dy * dy)))
        a = a if a >= 1 / 255 else 0.0
        np.testing.assert_allclose(rgb[i, j], a * s2[6:9, 0].astype(float), rtol=1e-6, atol=1e-9)


# ------------------------------------------------------------------ backward chain
@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_gradients_match_finite_differences_on_reference_scene(deg):
    """The reference's grad_check scene, exact gradient mode: every field of
    every splat vs FP64 central differences (the reference checks 5 % at SH 0)."""
    sc = gradcheck_scene(deg)
    n = sc["n"]
    fw = fp64_forward(sc["params"], n, deg, sc["cam"], sc["pv"], sc["W"], sc["H"],
                      sc["target"])
    bw = fp64_backward(fw, sc["params"], n, deg, sc["cam"], sc["pv"], sc["W"], sc["H"])
    fd = fd_grad(sc, pairs=fw["ta"])
    for k, name in enumerate(field_names(deg)):
        a, b = bw["gparams"][k * n:(k + 1) * n], fd[k * n:(k + 1) * n]
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-9, err_msg=name)


@pytest.mark.parametrize("seed,deg", [(0, 0), (1, 0), (2, 3)])
def test_gradients_match_finite_differences_on_random_scene(seed, deg):
    """General rotations, anisotropic scales and a perspective camera, so every
    quaternion component and covariance entry carries gradient."""
    sc = random_scene(seed, deg)
    n = sc["n"]
    fw = fp64_forward(sc["params"], n, deg, sc["cam"], sc["pv"], sc["W"], sc["H"],
                      sc["target"])
    assert fw["ta"]["n_isects"] > 2 * n
    bw = fp64_backward(fw, sc["params"], n, deg, sc["cam"], sc["pv"], sc["W"], sc["H"])
    fd = fd_grad(sc, pairs=fw["ta"])
    for k, name in enumerate(field_names(deg)):
        a, b = bw["gparams"][k * n:(k + 1) * n], fd[k * n:(k + 1) * n]
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-9, err_msg=name)


def test_reference_gradient_formula_differs_only_where_splats_overlap():
    """grad_mode 1 restates the reference's dL/dalpha (tile_rasterize_layer.cu:491-495).
    With one splat per pixel there is no colour behind and it equals the exact
    gradient; with overlap it does not, and only the exact one matches finite
    differences (previous test)."""
    sc = gradcheck_scene(0)
    n = sc["n"]
    args = (sc["params"], n, 0, sc["cam"], sc["pv"], sc["W"], sc["H"])
    fw = fp64_forward(*args, sc["target"])
    g0 = fp64_backward(fw, *args, grad_mode=0)["gparams"]
    g1 = fp64_backward(fw, *args, grad_mode=1)["gparams"]
    assert np.abs(g0 - g1).max() > 1e-4
    solo = dict(sc)
    p = sc["params"].copy()
    p.reshape(-1, n)[0] = np.arange(n) * 9.0 - 31.0  # spread out: no overlap
    p.reshape(-1, n)[1] = 0.0
    p.reshape(-1, n)[3:6] = 0.3
    solo["params"] = p
    solo["W"] = 80
    solo["target"] = np.full(80 * 32 * 3, 0.7)
    solo["pv"] = ortho(-40, 40, -16, 16, -4, 4)
    args = (p, n, 0, sc["cam"], solo["pv"], 80, 32)
    fw = fp64_forward(*args, solo["target"])
    g0 = fp64_backward(fw, *args, grad_mode=0)["gparams"]
    g1 = fp64_backward(fw, *args, grad_mode=1)["gparams"]
    np.testing.assert_allclose(g0, g1, rtol=1e-12, atol=1e-15)


def test_mse_closed_form():
    npix = 50
    rgb = np.full(3 * npix, 0.25)
    tgt = np.full(3 * npix, 0.75)
    loss, g = oracle.mse(rgb, tgt, npix)
    assert abs(loss - 3 * 0.25) < 1e-15
    np.testing.assert_allclose(g, 2 * (-0.5) / npix)


def test_adam_constant_gradient_closed_form():
    """With a constant gradient the bias-corrected moments equal g and g^2, so
    every step moves each parameter by lr * g / (|g| + eps)."""
    g = np.array([1e-3, -2.0, 0.5, 0.0])
    p = np.array([1.0, 2.0, 3.0, 4.0])
    m = np.zeros(4)
    v = np.zeros(4)
    lr = 0.01
    for t in range(1, 6):
        p_prev = p
        p, m, v = oracle.adam(p, g, m, v, lr, 0.9, 0.999, 1e-8, t)
        np.testing.assert_allclose(p_prev - p, lr * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_fp32_fit_step_composes_the_stages():
    """oracle.Fit (the CPU baseline) equals the stages chained by hand, and
    the loss falls."""
    import diffsplat_amd as ds

    W, H, n = 48, 32, 200
    params = ds.frustum_splats(3, n, aspect=W / H, ls_mean=-2.5)
    target = ds.target_image(3, W, H, 6)
    view, proj, cam = ds.fitter_camera(W, H)
    pv = ds.mat_mul(proj, view)
    lr6 = np.array([1e-2, 5e-3, 6e-4, 1e-2, 6e-3, 6e-3], np.float32)
    fit = oracle.Fit(params, target, n, W, H, 0, pv, cam, lr6)
    l0 = fit.step()
    p1 = fit.params()
    f32 = np.float32
    s3 = oracle.act_fwd(params, n, 0, cam, f32)
    s2 = oracle.proj_fwd(s3, n, pv, f32)
    ta = oracle.tile_assign_sort(s2, n, W, H)
    c = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H, f32)
    loss, dldc = oracle.mse(c["rgb"], target, W * H, f32)
    assert abs(loss - l0) <= 1e-7 * abs(l0)
    gs2, _ = oracle.comp_bwd(s2, n, ta["ids"], ta["ranges"], W, H, c["T"], c["last"], dldc, f32)
    gs3 = oracle.proj_bwd(s3, gs2, n, pv, f32)
    gp = oracle.act_bwd(params, s3, gs3, n, 0, cam, f32)
    lrs = np.repeat(lr6[[0, 1, 2, 3, 5]], [3 * n, 3 * n, 4 * n, 3 * n, n])
    want, _, _ = oracle.adam(params, gp, np.zeros_like(gp), np.zeros_like(gp), 0.0, 0.9, 0.999,
                             1e-8, 1, f32)
    want = params - lrs * gp / (np.abs(gp) + 1e-8)
    np.testing.assert_allclose(p1, want, rtol=1e-5, atol=1e-6)
    losses = [fit.step() for _ in range(30)]
    assert losses[-1] < 0.8 * l0


def _conic_g(s, j, px, py):
    """FP64 falloff exp(-d^2/2) of splat j at an NDC pixel centre."""
    Cm = np.array([[s[3, j], s[4, j]], [s[4, j], s[5, j]]])
    d = np.array([s[0, j] - px, s[1, j] - py])
    return math.exp(-0.5 * d @ np.linalg.solve(Cm, d))


def _grads(s2, n, W, H, ta, dldc):
    fw = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H)
    g, _, amb = oracle.comp_bwd(s2, n, ta["ids"], ta["ranges"], W, H, fw["T"], fw["last"], dldc,
                                px_margin=fw["margin"], amb_margin=5e-5)
    return g, amb, fw


@pytest.mark.parametrize("kind", ["threshold", "clamp", "termination"])
def test_flip_allowance_covers_decision_flips(kind):
    """The compositing backward's flip allowance (the bound the GPU parity
    tests grant elements near a discrete decision) against real flips: one
    splat's opacity is set so that a decision at one pixel lies 1e-7 (relative)
    on either side of its threshold; the gradients on the two sides differ by
    no more than the allowance computed on one side (plus 1e-4 relative for
    the continuous change of the opacity itself), and the flip changes some
    gradient by more than the strict 1e-4 relative tolerance."""
    s2, n, W, H = small_scene(n=150 if kind == "termination" else 30, seed=11)
    s = s2.reshape(10, n).astype(np.float64)
    if kind == "termination":
        s[9] = 0.95  # opaque: pixels stop at T <= 1e-4
    s2 = s.reshape(-1)
    ta = oracle.tile_assign_sort(s2.astype(np.float32), n, W, H)
    rng = np.random.default_rng(5)
    dldc = rng.normal(0, 1e-3, 3 * W * H)
    fw = oracle.comp_fwd(s2, n, ta["ids"], ta["ranges"], W, H)
    ntx = (W + 15) // 16
    ranges = ta["ranges"].reshape(-1, 2)
    # choose (pixel, splat): the pixel's last contributor for termination,
    # else the splat of the pixel's list whose falloff there suits the kind
    best = None
    for i in range(H):
        for jx in range(W):
            pix = i * W + jx
            t = (i // 16) * ntx + jx // 16
            px = float(np.float32(2) * (np.float32(jx) + np.float32(0.5)) / np.float32(W) - 1)
            py = float(1 - np.float32(2) * (np.float32(i) + np.float32(0.5)) / np.float32(H))
            if kind == "termination":
                if fw["T"][pix] > 1e-4:
                    continue
                k = int(fw["last"][pix])
                j = int(ta["ids"][k])
                g = _conic_g(s, j, px, py)
                a_old = min(0.99, s[9, j] * g)
                t_before = fw["T"][pix] / (1 - a_old)
                a_go = 1 - 1e-4 * (1 + 1e-7) / t_before    # T stays above: continue
                a_stop = 1 - 1e-4 * (1 - 1e-7) / t_before  # T drops below: stop
                if 0.02 < a_go < 0.98 and a_stop / g <= 1.0:
                    best = (pix, j, a_go / g, a_stop / g)
                    break
                continue
            for k in range(ranges[t, 0], ranges[t, 1]):
                if k > fw["last"][pix]:
                    break
                j = int(ta["ids"][k])
                g = _conic_g(s, j, px, py)
                if kind == "threshold" and 0.05 < g < 0.5:
                    best = (pix, j, (1 / 255) * (1 + 1e-7) / g, (1 / 255) * (1 - 1e-7) / g)
                elif kind == "clamp" and g > 0.992:
                    best = (pix, j, 0.99 * (1 - 1e-7) / g, 0.99 * (1 + 1e-7) / g)
                if best:
                    break
            if best:
                break
        if best:
            break
    assert best is not None
    pix, j, o_a, o_b = best
    sa, sb = s.copy(), s.copy()
    sa[9, j], sb[9, j] = o_a, o_b
    ga, amb, fwa = _grads(sa.reshape(-1), n, W, H, ta, dldc)
    gb, _, fwb = _grads(sb.reshape(-1), n, W, H, ta, dldc)
    assert fwa["margin"][pix] <= 5e-5  # the pixel is flagged ambiguous
    d = np.abs(ga - gb)
    assert np.all(d <= amb + 1e-4 * np.abs(ga) + 1e-15), float((d - amb).max())
    assert np.any(d > 1e-4 * np.abs(ga) + 1e-15)  # the flip is visible
