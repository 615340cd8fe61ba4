"""Error reporting on the GPU path. A sort look-back that gives up waiting
sets a device flag; that flag is sticky across forwards and graph replays and
is reported (and cleared) by the calls that synchronise with the host. The
give-up itself is forced from the host by writing the flag word, so no kernel
ever spins out."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import diffsplat_amd as ds
from diffsplat_amd._lib import RasterView

pytestmark = pytest.mark.gpu

DS_ERR_STATE = 5


def raise_flag(raster_handle):
    v = RasterView()
    ds.check(ds.lib().ds_raster_get_view(raster_handle, C.byref(v)), "view")
    one = np.ones(1, np.uint32)
    ds.check(ds.lib().ds_synchronize(), "sync")
    ds.check(ds.lib().ds_memcpy_htod(v.lookback_error, one.ctypes.data, 4), "htod")


def expect_state_error(fn):
    with pytest.raises(ds.DiffsplatError) as e:
        fn()
    assert e.value.status == DS_ERR_STATE
    assert "look-back" in str(e.value)


def test_fitter_reports_a_give_up_from_any_replayed_step():
    W = H = 64
    n = 400
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam)
    fit.set_target(ds.target_image(1, W, H, 8))
    fit.upload_params(ds.frustum_splats(1, n, aspect=1.0))
    fit.set_graph(True)
    fit.step(1)
    fit.synchronize()
    r = C.c_void_p()
    ds.check(ds.lib().ds_fitter_get_raster(fit.h, C.byref(r)), "raster")
    raise_flag(r)
    fit.step(20)  # 16-step graph replays and single steps: none clears the flag
    expect_state_error(fit.loss)
    fit.forward()
    assert np.isfinite(fit.loss())  # reported once, then cleared
    raise_flag(r)
    fit.step(3)
    expect_state_error(fit.synchronize)
    fit.synchronize()
    fit.close()


def test_multifit_and_renderer_report_a_give_up():
    sizes = [(64, 48, 300, 1), (40, 40, 200, 2)]
    inst = []
    for w, h, n, _ in sizes:
        view, proj, cam = ds.fitter_camera(w, h)
        inst.append((w, h, n, view, proj, cam))
    mf = ds.MultiFitter(inst)
    for k, (w, h, n, seed) in enumerate(sizes):
        mf.set_target(k, ds.target_image(seed, w, h, 8))
        mf.upload_params(k, ds.frustum_splats(seed, n, aspect=w / h))
    mf.set_graph(True)
    mf.step(2)
    mf.synchronize()
    r = C.c_void_p()
    ds.check(ds.lib().ds_multifit_get_raster(mf.h, C.byref(r)), "raster")
    raise_flag(r)
    mf.step(2)
    expect_state_error(mf.losses)
    assert np.all(np.isfinite(mf.losses()))
    mf.close()

    n = 500
    rend = ds.SplatRenderer(64, 64, n, 0)
    rend.upload_params(ds.frustum_splats(3, n, aspect=1.0))
    view, proj, cam = ds.fitter_camera(64, 64)
    rend.render(view, proj, cam)
    rend.synchronize()
    r = C.c_void_p()
    ds.check(ds.lib().ds_renderer_get_raster(rend.h, C.byref(r)), "raster")
    raise_flag(r)
    rend.render(view, proj, cam)
    expect_state_error(rend.synchronize)
    rend.synchronize()
    rend.close()
