"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/ds_api.h declares, validates arguments before touching a GPU, and its
host-side helpers (workload generators, camera matrices, defaults) behave."""
from __future__ import annotations

import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import diffsplat_amd as ds
from diffsplat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ds_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(ds_[a-z0-9_]+)\s*\(",
                                 text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    names = header_functions()
    assert len(names) > 60
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(header_functions()) == set(_lib.declared_symbols())


def test_library_reports_its_target():
    assert b"gfx950" in ds.lib().ds_version()
    assert ds.lib().ds_status_string(0) == b"DS_OK"
    assert ds.lib().ds_status_string(4) == b"DS_ERR_CAPACITY"


@pytest.mark.skipif(ds.has_gpu(), reason="checks the no-GPU path")
def test_no_gpu_here():
    assert ds.lib().ds_device_count() == 0


def test_invalid_arguments_are_rejected_before_any_device_call():
    L = ds.lib()
    h = C.c_void_p()
    assert L.ds_raster_create(None, C.byref(h)) == 1
    assert b"NULL" in L.ds_last_error()
    cfg = _lib.RasterConfig(64, 64, 10, 0, 0, 7)  # bad grad_mode
    assert L.ds_raster_create(C.byref(cfg), C.byref(h)) == 1
    assert b"grad_mode" in L.ds_last_error()
    assert L.ds_raster_set_tuning(None, 4, 4) == 1
    g = _lib.Gauss3D()
    g.count = 5
    s3 = _lib.Splat3D()
    cam = (C.c_float * 3)()
    assert L.ds_activ_fwd(C.byref(g), C.byref(s3), 0, cam, None) == 1  # NULL fields
    assert L.ds_activ_fwd(C.byref(g), C.byref(s3), 4, cam, None) == 1  # degree 4
    assert L.ds_project_bwd(None, None, None, None, 0, None) == 1
    grp = (_lib.AdamGroup * 1)()
    hp = _lib.AdamHParams(0.9, 0.999, 1e-8)
    assert L.ds_adam_step(grp, 99, C.byref(hp), 1, None) == 1
    assert L.ds_adam_step(grp, 0, C.byref(hp), 0, None) == 1  # step must be >= 1
    assert L.ds_sort_pairs_u64(None, None, None, None, 10, 65, None) == 1
    fc = _lib.FitterConfig()
    L.ds_fitter_default_config(32, 32, 4, C.byref(fc))
    fc.sh_degree = 5
    assert L.ds_fitter_create(C.byref(fc), C.byref(h)) == 1
    fake = (C.c_void_p * 2)(0x1000, 0x1000)  # never dereferenced: rejected first
    assert L.ds_batch_create(fake, 0, C.byref(h)) == 1
    assert L.ds_batch_create(fake, 2, C.byref(h)) == 1
    assert b"twice" in L.ds_last_error()
    nulls = (C.c_void_p * 1)(None)
    assert L.ds_batch_create(nulls, 1, C.byref(h)) == 1
    assert L.ds_batch_step(None, 1) == 1
    assert L.ds_renderer_autotune(None, 3, None, None) == 1
    # many fits as one, fused compositing, dispatch order
    inst = (_lib.RasterInstance * 2)(_lib.RasterInstance(64, 48, 10), _lib.RasterInstance(0, 8, 1))
    assert L.ds_raster_create_multi(inst, 0, 0, 0, 0, C.byref(h)) == 1
    assert L.ds_raster_create_multi(inst, 2, 0, 0, 0, C.byref(h)) == 1  # width 0
    assert b"size" in L.ds_last_error()
    assert L.ds_raster_create_multi(inst, 1, 0, 0, 5, C.byref(h)) == 1  # grad_mode
    assert L.ds_raster_splat_instances(None, None) == 1
    mfi = (_lib.MultiFitInstance * 1)()
    assert L.ds_multifit_create(mfi, 1, None, 0, C.byref(h)) == 1
    fc2 = _lib.FitterConfig()
    L.ds_fitter_default_config(32, 32, 4, C.byref(fc2))
    assert L.ds_multifit_create(mfi, 0, C.byref(fc2), 0, C.byref(h)) == 1
    mfi[0].width, mfi[0].height, mfi[0].count = 16, 16, -1
    assert L.ds_multifit_create(mfi, 1, C.byref(fc2), 0, C.byref(h)) == 1
    assert L.ds_multifit_step(None, 1) == 1 and L.ds_multifit_losses(None, None) == 1
    assert L.ds_raster_fwd_bwd(None, None, None, None, None, None, None, None) == 1
    assert L.ds_raster_composite_fwd_bwd(None, None, None, None, None, None, None, None) == 1
    assert L.ds_raster_set_tile_order(None, 1) == 1
    assert L.ds_activ_project_fwd_multi(None, None, 0, None, None, None) == 1
    with pytest.raises(ds.DiffsplatError):
        ds.check(L.ds_raster_create(None, C.byref(h)), "ds_raster_create")


def test_fitter_defaults_are_the_references():
    """image_fitter.cpp:18-34 camera, gaussian3d_adam.h:11-19 learning rates,
    adam.h:5-9 hyper-parameters."""
    fc = _lib.FitterConfig()
    ds.lib().ds_fitter_default_config(1280, 720, 60000, C.byref(fc))
    assert (fc.width, fc.height, fc.count, fc.sh_degree) == (1280, 720, 60000, 0)
    lr = [fc.lr_pos, fc.lr_scale, fc.lr_rot, fc.lr_sh_dc, fc.lr_sh_rest, fc.lr_opacity]
    np.testing.assert_allclose(lr, [1.6e-2, 1e-3, 6e-4, 6e-3, 6e-3, 6e-3], rtol=1e-6)
    np.testing.assert_allclose([fc.beta1, fc.beta2, fc.epsilon], [0.9, 0.999, 1e-8], rtol=1e-6)
    np.testing.assert_array_equal(np.array(fc.view), np.eye(4).reshape(-1))
    zr = 360 * 0.25
    want = np.array(ds.ortho(-640, 640, -360, 360, -zr, zr))
    np.testing.assert_allclose(np.array(fc.proj), want)
    assert fc.fused == 1 and fc.grad_mode == 0 and fc.lr_pos_decay == 1.0


def test_camera_matrices_follow_opengl_conventions():
    P = np.array(ds.perspective(math.radians(60), 16 / 9, 0.5, 20)).reshape(4, 4).T
    # a point on the near plane at the top edge of the view maps to ndc (0, 1, -1)
    t = math.tan(math.radians(30))
    c = P @ np.array([0, 0.5 * t, -0.5, 1.0])
    np.testing.assert_allclose(c[:3] / c[3], [0, 1, -1], atol=1e-6)
    c = P @ np.array([0, 0, -20.0, 1.0])
    np.testing.assert_allclose(c[2] / c[3], 1.0, atol=1e-6)
    O = np.array(ds.ortho(-2, 4, -1, 3, 0.5, 10)).reshape(4, 4).T
    np.testing.assert_allclose(O @ [4, 3, -10, 1], [1, 1, 1, 1], atol=1e-6)
    np.testing.assert_allclose(O @ [-2, -1, -0.5, 1], [-1, -1, -1, 1], atol=1e-6)
    V = np.array(ds.look_at([3, 1, 2], [0, 0, 0], [0, 1, 0])).reshape(4, 4).T
    np.testing.assert_allclose(V @ [3, 1, 2, 1], [0, 0, 0, 1], atol=1e-6)  # eye -> origin
    d = V @ np.array([0, 0, 0, 1.0])
    np.testing.assert_allclose(d[:2], [0, 0], atol=1e-6)  # target on the -z axis
    assert d[2] < 0
    np.testing.assert_allclose(np.linalg.det(V[:3, :3]), 1.0, atol=1e-6)
    A, B = np.random.default_rng(0).normal(size=(2, 16)).astype(np.float32)
    got = np.array(ds.mat_mul(A, B)).reshape(4, 4).T
    np.testing.assert_allclose(got, A.reshape(4, 4).T @ B.reshape(4, 4).T, rtol=1e-5, atol=1e-5)


def test_generators_are_deterministic_and_in_range():
    a = ds.target_image(1000, 64, 36, 12)
    b = ds.target_image(1000, 64, 36, 12)
    c = ds.target_image(1001, 64, 36, 12)
    assert a.shape == (64 * 36 * 3,) and np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() <= 1
    n = 500
    s = ds.frustum_splats(7, n, aspect=16 / 9)
    assert s.size == ds.slab_floats(n, 0) == 14 * n
    P = s.reshape(14, n)
    assert np.all((-P[2] >= 2) & (-P[2] <= 6))
    assert np.all(np.abs(P[0]) <= -P[2] * math.tan(math.radians(30)) * 16 / 9 + 1e-5)
    np.testing.assert_allclose(np.linalg.norm(P[6:10], axis=0), 1.0, rtol=1e-6)
    assert abs(P[3:6].mean() + 3.0) < 0.05 and abs(P[3:6].std() - 0.3) < 0.03
    assert np.all(P[13] == np.float32(0.1))
    rgb = 0.28209479177387814 * P[10:13] + 0.5
    assert rgb.min() >= -1e-6 and rgb.max() <= 1 + 1e-6
    q = ds.cluster_scene(3, 2000, 3, 64)
    assert q.size == ds.slab_floats(2000, 3) == (14 + 45) * 2000
    Q = q.reshape(-1, 2000)
    assert np.all(np.abs(Q[0:3]) <= 0.5)
    assert abs(Q[13:13 + 45].std() - 0.2) < 0.02
    f = ds.fitter_init(5, 300, 128, 96).reshape(14, 300)
    assert np.all(np.abs(f[0]) <= 64) and np.all(np.abs(f[1]) <= 48)
    assert np.all(f[6] == 1) and np.all(f[7:10] == 0)
    op = 1 / (1 + np.exp(-f[13]))
    assert op.min() >= 0.6 - 1e-6 and op.max() <= 1.0


def test_sh_band_counts():
    assert [ds.lib().ds_sh_bands(d) for d in range(4)] == [0, 3, 8, 15]
    assert ds.lib().ds_slab_floats(10, 3) == 590
