"""Runs each stage of the hot path on the GPU through the C-ABI and returns
numpy results, so parity tests can feed the oracle the same inputs."""
from __future__ import annotations

import ctypes as C

import numpy as np

import diffsplat_amd as ds


class Stages:
    """Device buffers for one scene: params -> Splat2D -> raster fwd/bwd."""

    def __init__(self, params, n, deg, W, H, view, proj, cam, grad_mode=0, isect_cap=0,
                 pix_per_thread=4):
        self.n, self.deg, self.W, self.H = n, deg, W, H
        self.grad_mode = grad_mode
        self.cam = np.ascontiguousarray(cam, np.float32)
        self.pv = ds.mat_mul(proj, view)
        self.params = np.ascontiguousarray(params, np.float32)
        self.d_params = ds.DeviceArray.from_numpy(self.params)
        self.d_s3 = ds.DeviceArray(4 * 13 * n)
        self.d_s2 = ds.DeviceArray(4 * 10 * n)
        self.d_gs2 = ds.DeviceArray(4 * 10 * n)
        self.d_gs3 = ds.DeviceArray(4 * 13 * n)
        self.d_gp = ds.DeviceArray(4 * ds.slab_floats(n, deg))
        self.d_rgb = ds.DeviceArray(4 * 3 * W * H)
        self.d_grad_rgb = ds.DeviceArray(4 * 3 * W * H)
        for d in (self.d_s3, self.d_s2, self.d_gs2, self.d_gs3, self.d_gp, self.d_rgb):
            d.zero()
        self.g = ds.gauss_view(self.d_params, n, deg)
        self.gg = ds.gauss_view(self.d_gp, n, deg)
        self.s3 = ds.splat3d_view(self.d_s3, n)
        self.gs3 = ds.splat3d_view(self.d_gs3, n)
        self.s2 = ds.splat2d_view(self.d_s2, n)
        self.gs2 = ds.splat2d_view(self.d_gs2, n)
        self.ras = ds.Rasterizer(W, H, n, isect_cap=isect_cap, keep_keys=True,
                                 grad_mode=grad_mode)
        ds.check(ds.lib().ds_raster_set_tuning(self.ras.h, pix_per_thread, pix_per_thread),
                 "set_tuning")

    # ---- forward
    def activ_fwd(self):
        ds.check(ds.lib().ds_activ_fwd(self.g, self.s3, self.deg, ds.fptr(self.cam), None),
                 "ds_activ_fwd")
        return self.d_s3.download(np.float32, 13 * self.n)

    def project_fwd(self):
        ds.check(ds.lib().ds_project_fwd(self.s3, self.s2, ds.fptr(self.pv), None),
                 "ds_project_fwd")
        return self.d_s2.download(np.float32, 10 * self.n)

    def activ_project_fwd(self):
        ds.check(ds.lib().ds_activ_project_fwd(self.g, self.s2, self.deg, ds.fptr(self.cam),
                                               ds.fptr(self.pv), None), "ds_activ_project_fwd")
        return self.d_s2.download(np.float32, 10 * self.n)

    def upload_s2(self, s2):
        self.d_s2.upload(np.ascontiguousarray(s2, np.float32))

    def raster_fwd(self, target=None):
        if target is not None:
            d_t = ds.DeviceArray.from_numpy(np.ascontiguousarray(target, np.float32))
            d_loss = ds.DeviceArray(4)
            self.ras.forward(self.s2, self.d_rgb, d_t, d_loss, self.d_grad_rgb)
            st = self.ras.state()
            st["loss"] = float(d_loss.download(np.float32, 1)[0])
            st["grad_rgb"] = self.d_grad_rgb.download(np.float32, 3 * self.W * self.H)
        else:
            self.ras.forward(self.s2, self.d_rgb)
            st = self.ras.state()
        st["rgb"] = self.d_rgb.download(np.float32, 3 * self.W * self.H)
        return st

    # ---- backward
    def raster_bwd(self, grad_rgb):
        self.d_grad_rgb.upload(np.ascontiguousarray(grad_rgb, np.float32))
        ds.check(ds.lib().ds_splat2d_zero(self.gs2, None), "ds_splat2d_zero")
        self.ras.backward(self.s2, self.d_grad_rgb, self.gs2)
        ds.check(ds.lib().ds_synchronize(), "sync")
        return self.d_gs2.download(np.float32, 10 * self.n)

    def upload_gs2(self, gs2):
        self.d_gs2.upload(np.ascontiguousarray(gs2, np.float32))

    def project_bwd(self):
        ds.check(ds.lib().ds_project_bwd(self.s3, self.gs2, self.gs3, ds.fptr(self.pv),
                                         self.grad_mode, None),
                 "ds_project_bwd")
        return self.d_gs3.download(np.float32, 13 * self.n)

    def activ_bwd(self):
        ds.check(ds.lib().ds_activ_bwd(self.g, self.s3, self.gs3, self.gg, self.deg,
                                       ds.fptr(self.cam), self.grad_mode, None), "ds_activ_bwd")
        return self.d_gp.download(np.float32, ds.slab_floats(self.n, self.deg))

    def activ_project_bwd(self):
        ds.check(ds.lib().ds_activ_project_bwd(self.g, self.gs2, self.gg, self.deg,
                                               ds.fptr(self.cam), ds.fptr(self.pv),
                                               self.grad_mode, None),
                 "ds_activ_project_bwd")
        return self.d_gp.download(np.float32, ds.slab_floats(self.n, self.deg))


def frustum_scene(seed, n, W, H, ls_mean=-3.0):
    params = ds.frustum_splats(seed, n, aspect=W / H, ls_mean=ls_mean)
    view, proj, cam = ds.fitter_camera(W, H)
    return params, view, proj, cam


def with_sh(params, n, deg, seed):
    """Adds random SH rest coefficients to a degree-0 slab."""
    nb = ds.sh_bands(deg)
    rng = np.random.default_rng(seed)
    out = np.concatenate([params[:13 * n], rng.normal(0, 0.2, 3 * nb * n).astype(np.float32),
                          params[13 * n:]])
    return out.astype(np.float32)
