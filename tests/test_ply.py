"""3DGS PLY scene IO and scene normalisation (host code, CPU tests).

The files are written and read here by an independent numpy implementation
of the format PLYLoader/PLYSaver use (src/io/ply_io.cpp): binary
little-endian float vertex properties by name, f_rest channel-first, OpenCV
axes in the file. normalize_scene is checked bit for bit against a
restatement of SplatUtils::normalizeScene (src/utils/splat_utils.h:53-93)
in sequential float32 arithmetic."""
from __future__ import annotations

import math

import numpy as np
import pytest

import diffsplat_amd as ds

BANDS = {0: 0, 1: 3, 2: 8, 3: 15}


def write_ply(path, props: dict, order, extra_header=""):
    """props: name -> float32 array (all the same length), written in `order`."""
    n = len(next(iter(props.values())))
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    head += [f"property float {k}" for k in order]
    if extra_header:
        head.append(extra_header)
    head.append("end_header")
    data = np.stack([np.asarray(props[k], np.float32) for k in order], axis=1)
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(data.astype("<f4").tobytes())


def read_ply(path):
    raw = open(path, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    lines = raw[:end].decode().splitlines()
    n = int([l for l in lines if l.startswith("element vertex")][0].split()[2])
    names = [l.split()[2] for l in lines if l.startswith("property float")]
    data = np.frombuffer(raw[end:], "<f4").reshape(n, len(names))
    return names, {k: data[:, i] for i, k in enumerate(names)}


def random_props(n, deg, seed):
    rng = np.random.default_rng(seed)
    p = {k: rng.normal(0, 1, n).astype(np.float32) for k in
         ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
          "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]}
    for k in range(3 * BANDS[deg]):
        p[f"f_rest_{k}"] = rng.normal(0, 1, n).astype(np.float32)
    return p


def expected_slab(p, n, deg):
    """Slab layout: pos xyz, scale xyz, rot wxyz, dc rgb, sh_rest r/g/b band-major, logit;
    OpenCV -> OpenGL flips y, z and the quaternion's y, z."""
    nb = BANDS[deg]
    fields = [p["x"], -p["y"], -p["z"], p["scale_0"], p["scale_1"], p["scale_2"],
              p["rot_0"], p["rot_1"], -p["rot_2"], -p["rot_3"],
              p["f_dc_0"], p["f_dc_1"], p["f_dc_2"]]
    for c in range(3):
        for b in range(nb):
            fields.append(p[f"f_rest_{c * nb + b}"])
    fields.append(p["opacity"])
    return np.concatenate(fields).astype(np.float32)


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_load_3dgs_file_with_normals_in_any_order(tmp_path, deg):
    n = 37
    p = random_props(n, deg, deg)
    order = list(p)[::-1]  # reversed property order, normals included
    path = tmp_path / "scene.ply"
    write_ply(path, p, order)
    slab, count, d = ds.load_ply(path)
    assert (count, d) == (n, deg)
    np.testing.assert_array_equal(slab, expected_slab(p, n, deg))


def test_degree_detection_follows_the_f_rest_count(tmp_path):
    # 3 f_rest properties per channel-first band: >= 45 -> 3, >= 24 -> 2, >= 9 -> 1
    for count, deg in [(0, 0), (6, 0), (9, 1), (20, 1), (24, 2), (44, 2), (45, 3), (48, 3)]:
        p = random_props(5, 0, count)
        for k in range(count):
            p[f"f_rest_{k}"] = np.full(5, k, np.float32)
        path = tmp_path / f"d{count}.ply"
        write_ply(path, p, list(p))
        slab, n, d = ds.load_ply(path)
        assert d == deg, (count, d)
        nb = BANDS[deg]
        rest = slab[13 * 5:13 * 5 + 3 * nb * 5].reshape(3, nb, 5)
        for c in range(3):
            for b in range(nb):
                assert np.all(rest[c, b] == c * nb + b)


@pytest.mark.parametrize("deg", [0, 3])
def test_save_writes_the_reference_layout_and_round_trips(tmp_path, deg):
    n = 50
    slab = np.random.default_rng(4).normal(0, 1, ds.slab_floats(n, deg)).astype(np.float32)
    path = tmp_path / "out.ply"
    ds.save_ply(path, slab, n, deg)
    names, cols = read_ply(path)
    nb = BANDS[deg]
    want = (["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"] +
            [f"f_rest_{k}" for k in range(3 * nb)] +
            ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
    assert names == want
    np.testing.assert_array_equal(expected_slab(cols, n, deg), slab)
    back, n2, d2 = ds.load_ply(path)
    assert (n2, d2) == (n, deg)
    np.testing.assert_array_equal(back, slab)


def normalize_reference(slab, n, scale):
    """normalizeScene in sequential float32 arithmetic."""
    s = slab.copy()
    f = np.float32
    cx = cy = cz = f(0)
    for i in range(n):
        cx = f(cx + s[i])
        cy = f(cy + s[n + i])
        cz = f(cz + s[2 * n + i])
    inv = f(f(1) / f(n))
    cx, cy, cz = f(cx * inv), f(cy * inv), f(cz * inv)
    s[0:n] = s[0:n] - cx
    s[n:2 * n] = s[n:2 * n] - cy
    s[2 * n:3 * n] = s[2 * n:3 * n] - cz
    ext = f(np.abs(s[0:3 * n]).max())
    if ext > 0:
        lin = f(f(scale) / ext)
        lg = f(math.log(lin))
        s[0:3 * n] = s[0:3 * n] * lin
        s[3 * n:6 * n] = s[3 * n:6 * n] + lg
    return s


def test_normalize_scene_is_bit_identical_to_the_reference_order():
    n = 777
    slab = ds.cluster_scene(11, n, 1, 8)
    slab[0:3 * n] = slab[0:3 * n] * np.float32(37.0) + np.float32(5.0)
    want = normalize_reference(slab, n, 0.8)
    got = ds.normalize_scene(slab.copy(), n, 0.8)
    np.testing.assert_array_equal(got, want)
    assert abs(np.abs(got[0:3 * n]).max() - 0.8) <= 1e-6
    np.testing.assert_array_equal(got[6 * n:], slab[6 * n:])  # rotations, colours untouched
    # What the float64 mathematics fixes: centred positions, one uniform scale
    # factor 0.8 / extent, and log-scales shifted by its log.
    pos = slab[0:3 * n].astype(np.float64).reshape(3, n)
    centred = pos - pos.mean(axis=1, keepdims=True)
    k = 0.8 / np.abs(centred).max()
    np.testing.assert_allclose(got[0:3 * n].reshape(3, n), centred * k, atol=2e-6)
    np.testing.assert_allclose(got[3 * n:6 * n] - slab[3 * n:6 * n], math.log(k), atol=2e-6)


def test_rejects_what_the_reference_would_misread(tmp_path):
    p = random_props(3, 0, 1)
    path = tmp_path / "scene.ply"
    write_ply(path, p, list(p))
    with pytest.raises(ds.DiffsplatError, match="extension"):
        ds.load_ply(str(path)[:-4] + ".obj")
    bad = tmp_path / "ascii.ply"
    bad.write_bytes(open(path, "rb").read().replace(b"binary_little_endian", b"ascii"))
    with pytest.raises(ds.DiffsplatError, match="binary_little_endian"):
        ds.load_ply(bad)
    missing = tmp_path / "missing.ply"
    q = dict(p)
    del q["opacity"]
    write_ply(missing, q, list(q))
    with pytest.raises(ds.DiffsplatError, match="opacity"):
        ds.load_ply(missing)
    short = tmp_path / "short.ply"
    short.write_bytes(open(path, "rb").read()[:-8])
    with pytest.raises(ds.DiffsplatError, match="ends early"):
        ds.load_ply(short)
    uchar = tmp_path / "uchar.ply"
    raw = open(path, "rb").read().replace(b"property float nx", b"property uchar nx")
    uchar.write_bytes(raw)
    with pytest.raises(ds.DiffsplatError, match="not float"):
        ds.load_ply(uchar)
