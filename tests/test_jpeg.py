"""JPEG target images (host/jpeg.cpp), on the CPU.

The reference decodes JPEG with stb_image (src/io/image_io.cpp:26) and
resizes with stb_image_resize2 (:44). Pins:
  - exact invariants of the entropy decoding: a progressive file and a
    baseline file of the same image, quality and subsampling decode to the
    same pixels (the same quantised coefficients), and so do files with and
    without restart markers;
  - against Pillow's decoder (libjpeg-turbo): grey images to 1 of 255 (the
    inverse DCTs' rounding), colour within 3 and mostly within 1 (the
    decoders' integer upsampling and colour transforms round differently;
    stb_image and Pillow differ from each other by as much);
  - against the reference's own stb_image and stb_image_resize2, compiled
    from the reference tree into oracle/_ref by `make -C oracle ref`, when
    that library is present: its five JPEG test images (where the reference
    tree is present), and the loader's resize of a decoded image."""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np
import pytest
from PIL import Image

import diffsplat_amd as ds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libstbref.so")
REF_IMAGES = sorted(glob.glob("/root/reference/data/img/*.jpg"))


def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference tree)")
    lib = C.CDLL(REF_SO)
    lib.ref_load_rgb.restype = C.POINTER(C.c_ubyte)
    lib.ref_load_rgb.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.ref_free.argtypes = [C.c_void_p]
    lib.ref_resize_rgb.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
    return lib


def stb_decode(path):
    lib = ref_lib()
    w, h = C.c_int(), C.c_int()
    p = lib.ref_load_rgb(str(path).encode(), C.byref(w), C.byref(h))
    assert p, path
    a = np.ctypeslib.as_array(p, (h.value, w.value, 3)).copy()
    lib.ref_free(p)
    return a


def decode(path):
    """The engine's decode at the image's own size (the loader's resize is
    then the identity), as 8-bit values."""
    w, h = C.c_int32(), C.c_int32()
    ds.check(ds.lib().ds_image_info(str(path).encode(), C.byref(w), C.byref(h)), "info")
    img, _, _ = ds.load_image(path, w.value, h.value)
    return np.round(img.reshape(h.value, w.value, 3) * 255).astype(np.int32)


def picture(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([128 + 90 * np.sin(x / (5 + 3 * c) + c) * np.cos(y / (4 + c) - c)
                    for c in range(3)], -1)
    img += rng.normal(0, 12, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


def write_jpeg(tmp_path, name, img, **kw):
    p = tmp_path / name
    mode = "L" if img.ndim == 2 else "RGB"
    Image.fromarray(img, mode).save(p, "JPEG", **kw)
    return p


CASES = [  # (height, width, subsampling (0 = 4:4:4, 1 = 4:2:2, 2 = 4:2:0), quality)
    (64, 80, 2, 90), (37, 23, 2, 75), (48, 48, 1, 60), (33, 71, 0, 95), (120, 90, 2, 40)]


@pytest.mark.parametrize("h,w,sub,q", CASES)
def test_progressive_and_restarts_decode_to_the_baseline_pixels(tmp_path, h, w, sub, q):
    img = picture(h, w, h * w)
    base = decode(write_jpeg(tmp_path, "b.jpg", img, quality=q, subsampling=sub))
    prog = decode(write_jpeg(tmp_path, "p.jpg", img, quality=q, subsampling=sub,
                             progressive=True))
    rst = decode(write_jpeg(tmp_path, "r.jpg", img, quality=q, subsampling=sub,
                            restart_marker_blocks=3))
    assert np.array_equal(prog, base)
    assert np.array_equal(rst, base)


@pytest.mark.parametrize("h,w,q", [(64, 64, 90), (29, 41, 50), (100, 72, 98)])
def test_grey_matches_pillow_to_one_level(tmp_path, h, w, q):
    img = picture(h, w, 3)[..., 0]
    p = write_jpeg(tmp_path, "g.jpg", img, quality=q, progressive=(q == 50))
    got = decode(p)
    want = np.asarray(Image.open(p).convert("L")).astype(np.int32)
    assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    d = np.abs(got[..., 0] - want)
    print(f"[jpeg] grey {w}x{h} q{q}: max {d.max()} mean {d.mean():.4f}")
    assert d.max() <= 1


@pytest.mark.parametrize("h,w,sub,q", CASES)
def test_colour_matches_pillow(tmp_path, h, w, sub, q):
    img = picture(h, w, 11)
    p = write_jpeg(tmp_path, "c.jpg", img, quality=q, subsampling=sub)
    got = decode(p)
    want = np.asarray(Image.open(p).convert("RGB")).astype(np.int32)
    d = np.abs(got - want)
    print(f"[jpeg] colour {w}x{h} sub{sub} q{q}: max {d.max()} mean {d.mean():.4f} "
          f"over1 {(d > 1).mean():.5f}")
    # libjpeg's 2x horizontal-only upsampling alternates its rounding bias
    assert d.max() <= 3 and (d > 1).mean() <= (0.06 if sub == 1 else 0.02)


@pytest.mark.parametrize("h,w,sub,q", CASES)
def test_colour_matches_the_references_stb_image(tmp_path, h, w, sub, q):
    img = picture(h, w, 12)
    p = write_jpeg(tmp_path, "s.jpg", img, quality=q, subsampling=sub, progressive=(q < 70))
    got = decode(p)
    d = np.abs(got - stb_decode(p).astype(np.int32))
    if sub == 1:
        # stb_image's 2x horizontal-only upsampler weights the last two chroma
        # samples the wrong way round for its second-to-last output column
        # (Pillow agrees with this decoder there); that column is left out
        d = np.delete(d, 2 * ((w + 1) // 2) - 2, axis=1)
    print(f"[jpeg] vs stb {w}x{h} sub{sub} q{q}: max {d.max()} mean {d.mean():.4f} "
          f"over1 {(d > 1).mean():.5f}")
    assert d.max() <= 3 and (d > 1).mean() <= 0.02


@pytest.mark.skipif(not REF_IMAGES, reason="the reference tree is not present")
@pytest.mark.parametrize("path", REF_IMAGES, ids=[os.path.basename(p) for p in REF_IMAGES])
def test_reference_test_images_match_stb_image(path):
    """The reference's own data/img JPEGs (one baseline, four progressive,
    all 4:2:0)."""
    got = decode(path)
    want = stb_decode(path).astype(np.int32)
    assert got.shape == want.shape
    d = np.abs(got - want)
    print(f"[jpeg] {os.path.basename(path)}: max {d.max()} mean {d.mean():.4f} "
          f"over1 {(d > 1).mean():.5f}")
    assert d.max() <= 3 and (d > 1).mean() <= 0.01


@pytest.mark.parametrize("sw,sh,W,H,pad", [(100, 80, 64, 64, 0), (64, 48, 200, 150, 0),
                                           (301, 203, 256, 256, 3), (640, 360, 256, 256, 0)])
def test_loader_resize_matches_stb_image_resize2(tmp_path, sw, sh, W, H, pad):
    lib = ref_lib()
    img = picture(sh, sw, 5)
    p = tmp_path / "r.ppm"
    p.write_bytes(b"P6\n%d %d\n255\n" % (sw, sh) + img.tobytes())
    got, _, _ = ds.load_image(p, W, H, pad)
    got = np.round(got.reshape(H, W, 3) * 255).astype(np.int32)
    scale = min(np.float32(W - 2 * pad) / np.float32(sw), np.float32(H - 2 * pad) / np.float32(sh))
    fw, fh = int(sw * scale), int(sh * scale)
    dst = np.zeros((fh, fw, 3), np.uint8)
    assert lib.ref_resize_rgb(img.ctypes.data, sw, sh, dst.ctypes.data, fw, fh)
    ox, oy = (W - fw) // 2, (H - fh) // 2
    want = np.zeros((H, W, 3), np.int32)
    want[oy:oy + fh, ox:ox + fw] = dst
    assert np.abs(got - want).max() <= 1


def test_rejects_arithmetic_coding_and_truncation(tmp_path):
    img = picture(16, 16, 1)
    good = write_jpeg(tmp_path, "a.jpg", img).read_bytes()
    bad = bytearray(good)
    i = bad.index(b"\xff\xc0")
    bad[i + 1] = 0xC9  # SOF9: arithmetic coding
    p = tmp_path / "arith.jpg"
    p.write_bytes(bytes(bad))
    with pytest.raises(ds.DiffsplatError, match="unsupported JPEG process"):
        ds.load_image(p, 16, 16)
    p.write_bytes(good[:len(good) // 3])
    with pytest.raises(ds.DiffsplatError):
        ds.load_image(p, 16, 16)
