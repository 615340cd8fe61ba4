"""Pins the oracle's adaptive density control (clone / split / prune, Kerbl et
al., SIGGRAPH 2023, sec. 5.2; a TODO in the reference, README.md:10) without
the GPU:
  - the actions against an independent numpy statement of the rules;
  - survivors and clones as exact copies, with their Adam moments (clones and
    split samples with zero moments), in the output order survivors, then new
    splats in splat order;
  - split samples against numpy's own hash-driven Box-Muller draw and
    scipy's rotation, and their distribution against N(mu, R S^2 R^T) over
    many seeds (the covariance a dropped rotation or scale would break);
  - split scales divided by 1.6."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle

M64 = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform(seed, i, s, k):
    h = mix64(seed ^ mix64((i << 8) | (s << 2) | k))
    return ((h >> 40) + 1) / 16777216.0


def scene(n=400, seed=0, deg=0):
    rng = np.random.default_rng(seed)
    nb = oracle.bands(deg)
    nf = 14 + 3 * nb
    slab = np.zeros((nf, n), np.float32)
    slab[0:3] = rng.normal(0, 1, (3, n))
    slab[3:6] = rng.normal(-3.0, 0.8, (3, n))
    slab[6:10] = rng.normal(0, 1, (4, n))
    slab[10:13] = rng.normal(0, 0.5, (3, n))
    if nb:
        slab[13:13 + 3 * nb] = rng.normal(0, 0.2, (3 * nb, n))
    slab[-1] = rng.normal(-1.0, 2.5, n)  # some below logit(0.005) = -5.29
    m = rng.normal(0, 1e-3, (nf, n)).astype(np.float32)
    v = rng.uniform(0, 1e-6, (nf, n)).astype(np.float32)
    accum = rng.uniform(0, 4e-3, n).astype(np.float32)
    visits = rng.integers(0, 12, n).astype(np.uint32)
    return slab.reshape(-1), m.reshape(-1), v.reshape(-1), accum, visits, nf


def numpy_actions(slab, n, nf, accum, visits, p):
    S = slab.reshape(nf, n)
    g = np.where(visits > 0, accum / np.maximum(visits, 1).astype(np.float32), np.float32(0))
    ls = S[3:6].max(axis=0)
    faint = S[-1] < np.float32(p["min_logit"])
    hot = g >= np.float32(p["grad_threshold"])
    small = ls <= np.float32(p["clone_max_log_scale"])
    big = np.zeros(n, bool)
    if p["prune_big"]:
        big = ls > np.float32(p["prune_max_log_scale"])
    a = np.where(big, 0, 1)
    a = np.where(hot & small, 2, a)
    a = np.where(hot & ~small, 3, a)
    a = np.where(faint, 0, a)
    return a.astype(np.uint8)


@pytest.mark.parametrize("deg,prune_frac", [(0, 0.0), (0, 0.08), (2, 0.0)])
def test_actions_copies_moments_and_order(deg, prune_frac):
    slab, m, v, accum, visits, nf = scene(deg=deg, seed=deg + 1)
    n = len(accum)
    p = oracle.densify_params(grad_threshold=2e-4, percent_dense=0.05, scene_extent=1.0,
                              max_scale_frac=prune_frac, seed=7)
    r = oracle.densify(slab, m, v, n, deg, accum, visits, p)
    act = r["action"]
    want = numpy_actions(slab, n, nf, accum, visits, p)
    assert np.array_equal(act, want)
    assert {0, 1, 2, 3} <= set(act.tolist()) or prune_frac == 0.0
    S, Mm, Vv = slab.reshape(nf, n), m.reshape(nf, n), v.reshape(nf, n)
    ls = S[3:6].max(axis=0)
    child_big = (ls - np.float32(p["split_log_div"]) > np.float32(p["prune_max_log_scale"])) \
        if p["prune_big"] else np.zeros(n, bool)
    survivors = np.where((act == 1) | (act == 2))[0]
    clones = np.where(act == 2)[0]
    splits = np.where((act == 3) & ~child_big)[0]
    n_out = len(survivors) + len(clones) + 2 * len(splits)
    assert r["n_out"] == n_out
    O = r["params"].reshape(nf, n_out)
    Om, Ov = r["m"].reshape(nf, n_out), r["v"].reshape(nf, n_out)
    k = len(survivors)
    assert np.array_equal(O[:, :k], S[:, survivors].astype(np.float64))
    assert np.array_equal(Om[:, :k], Mm[:, survivors]) and np.array_equal(Ov[:, :k], Vv[:, survivors])
    # new splats in splat order: a clone adds one copy, a split two samples
    new_src = np.sort(np.concatenate([clones, splits, splits]), kind="stable")
    assert np.all(Om[:, k:] == 0) and np.all(Ov[:, k:] == 0)
    for j, i in enumerate(new_src):
        col = O[:, k + j]
        if act[i] == 2:
            assert np.array_equal(col, S[:, i].astype(np.float64))
        else:
            assert np.array_equal(col[6:], S[6:, i].astype(np.float64))
            np.testing.assert_array_equal(col[3:6], (S[3:6, i] - np.float32(p["split_log_div"]))
                                          .astype(np.float64))


def test_split_samples_follow_the_hashed_box_muller_draw():
    slab, m, v, accum, visits, nf = scene(n=200, seed=3)
    n = 200
    accum[:] = 1.0
    visits[:] = 1
    p = oracle.densify_params(grad_threshold=1e-3, percent_dense=1e-6, scene_extent=1.0,
                              min_opacity=1e-9, seed=12345)
    r = oracle.densify(slab, m, v, n, 0, accum, visits, p)
    assert np.all(r["action"] == 3) and r["n_out"] == 2 * n
    S = slab.reshape(nf, n).astype(np.float64)
    O = r["params"].reshape(nf, 2 * n)
    for i in range(0, n, 17):
        R = Rotation.from_quat([S[7, i], S[8, i], S[9, i], S[6, i]]).as_matrix()
        sc = np.exp(S[3:6, i])
        for s in range(2):
            u = [uniform(12345, i, s, k) for k in range(4)]
            r0, r1 = math.sqrt(-2 * math.log(u[0])), math.sqrt(-2 * math.log(u[2]))
            z = np.array([r0 * math.cos(2 * math.pi * u[1]), r0 * math.sin(2 * math.pi * u[1]),
                          r1 * math.cos(2 * math.pi * u[3])])
            np.testing.assert_allclose(O[0:3, 2 * i + s], S[0:3, i] + R @ (sc * z), rtol=0,
                                       atol=1e-12)


def test_split_samples_have_the_splats_covariance():
    """Over 4000 seeds the sample offsets of one anisotropic, rotated splat
    have mean 0 and covariance R S^2 R^T."""
    n = 1
    slab = np.zeros(14, np.float32)
    slab[3:6] = np.log([0.3, 0.1, 0.05])
    q = np.array([0.9, 0.3, -0.2, 0.25])
    slab[6:10] = q
    slab[13] = 2.0
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
    cov = R @ np.diag(np.exp(2 * slab[3:6].astype(np.float64))) @ R.T
    offs = []
    for seed in range(4000):
        p = oracle.densify_params(grad_threshold=0.0, percent_dense=1e-6, scene_extent=1.0,
                                  seed=seed)
        r = oracle.densify(slab, np.zeros(14, np.float32), np.zeros(14, np.float32), n, 0,
                           np.ones(1, np.float32), np.ones(1, np.uint32), p)
        offs.append(r["params"].reshape(14, 2)[0:3].T)
    d = np.concatenate(offs)
    assert np.abs(d.mean(axis=0)).max() < 4 * 0.3 / math.sqrt(len(d))
    emp = np.cov(d.T)
    assert np.abs(emp - cov).max() < 0.08 * np.abs(cov).max()
