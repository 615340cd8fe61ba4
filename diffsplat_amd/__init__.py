"""diffsplat_amd -- differentiable 3D Gaussian splatting for MI355X (gfx950).

The compute path is libdiffsplat.so (hand-written HIP kernels behind the C-ABI
in include/ds_api.h, with the reference's layer classes in C++ on top). This
package is a thin Python handle on that library: synthetic workloads, camera
matrices and the fitter / renderer / rasterizer objects.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import (AdamGroup, AdamHParams, CameraInput, CameraSpeeds, DeviceArray,
                   DensifyConfig, DensifyResult, DiffsplatError, FitterConfig, Gauss3D,
                   MultiFitInstance,
                   RasterConfig, RasterInstance, RasterView, check, download_ptr, farr, fptr,
                   Splat2D, gauss_view, has_gpu, lib, splat2d_view, splat3d_view)

__all__ = [
    "lib", "has_gpu", "check", "DiffsplatError", "DeviceArray", "gauss_view", "splat3d_view",
    "splat2d_view", "sh_bands", "slab_floats", "target_image", "frustum_splats",
    "cluster_scene", "fitter_init", "perspective", "ortho", "look_at", "mat_mul",
    "fitter_camera", "orbit_cameras", "ImageFitter", "SplatRenderer", "Rasterizer",
    "MultiRasterizer", "MultiFitter",
    "psnr_from_loss", "Input", "ArcballCamera", "FlyCamera",
]

SH_C0 = 0.28209479177387814


def sh_bands(deg: int) -> int:
    return (0, 3, 8, 15)[max(0, min(3, deg))]


def slab_floats(count: int, deg: int) -> int:
    return (14 + 3 * sh_bands(deg)) * count


# ---------------------------------------------------------------- workloads
def target_image(seed: int, width: int, height: int, n_ellipses: int) -> np.ndarray:
    out = np.empty(width * height * 3, np.float32)
    lib().ds_gen_target_image(seed, width, height, n_ellipses, fptr(out))
    return out


def frustum_splats(seed: int, count: int, fovy_deg: float = 60.0, aspect: float = 1.0,
                   zmin: float = 2.0, zmax: float = 6.0, ls_mean: float = -3.0,
                   ls_std: float = 0.3, logit_opacity: float = 0.1) -> np.ndarray:
    out = np.empty(slab_floats(count, 0), np.float32)
    lib().ds_gen_frustum_splats(seed, count, fovy_deg, aspect, zmin, zmax, ls_mean, ls_std,
                                logit_opacity, fptr(out))
    return out


def cluster_scene(seed: int, count: int, sh_degree: int = 3, n_clusters: int = 64,
                  ls_mean: float = -4.5, ls_std: float = 0.5) -> np.ndarray:
    out = np.empty(slab_floats(count, sh_degree), np.float32)
    lib().ds_gen_cluster_scene(seed, count, sh_degree, n_clusters, ls_mean, ls_std, fptr(out))
    return out


def fitter_init(seed: int, count: int, width: int, height: int) -> np.ndarray:
    out = np.empty(slab_floats(count, 0), np.float32)
    lib().ds_gen_fitter_init(seed, count, width, height, fptr(out))
    return out


# ---------------------------------------------------------------- cameras
def perspective(fovy_rad: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    m = np.empty(16, np.float32)
    lib().ds_mat_perspective(fovy_rad, aspect, znear, zfar, fptr(m))
    return m


def ortho(l, r, b, t, n, f) -> np.ndarray:
    m = np.empty(16, np.float32)
    lib().ds_mat_ortho(l, r, b, t, n, f, fptr(m))
    return m


def look_at(eye, center, up) -> np.ndarray:
    m = np.empty(16, np.float32)
    e, c, u = (np.asarray(v, np.float32) for v in (eye, center, up))
    lib().ds_mat_look_at(fptr(e), fptr(c), fptr(u), fptr(m))
    return m


def mat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    m = np.empty(16, np.float32)
    lib().ds_mat_mul(fptr(np.ascontiguousarray(a, np.float32)),
                     fptr(np.ascontiguousarray(b, np.float32)), fptr(m))
    return m


def fitter_camera(width: int, height: int, fovy_deg: float = 60.0):
    """Pinhole camera at the origin looking down -z (the frustum workloads)."""
    view = np.eye(4, dtype=np.float32).reshape(-1)
    proj = perspective(math.radians(fovy_deg), width / height, 0.5, 20.0)
    return view, proj, np.zeros(3, np.float32)


def orbit_cameras(n: int, width: int, height: int, radius: float = 3.0,
                  fovy_deg: float = 50.0, elevation: float = 0.35):
    """n look-at cameras on a circle of `radius` around the origin."""
    proj = perspective(math.radians(fovy_deg), width / height, 0.1, 20.0)
    cams = []
    for k in range(n):
        a = 2.0 * math.pi * k / n
        eye = np.array([radius * math.cos(a), radius * elevation, radius * math.sin(a)],
                       np.float32)
        cams.append((look_at(eye, (0, 0, 0), (0, 1, 0)), proj, eye))
    return cams


KEY_BITS = {"shift": 1 << 0, "ctrl": 1 << 1, "w": 1 << 2, "a": 1 << 3, "s": 1 << 4,
            "d": 1 << 5, "r": 1 << 6, "f": 1 << 7, "q": 1 << 8, "e": 1 << 9}


class Input:
    """One frame of user input (src/input/input.h:12-16): cursor movement in
    pixels (screen y down), wheel delta, left button, and the held keys by
    name ("shift", "ctrl", "w", "a", "s", "d", "r", "f", "q", "e")."""

    def __init__(self, mouse_delta=(0.0, 0.0), scroll_delta: float = 0.0,
                 mouse_left_held: bool = False, keys=()):
        self.mouse_delta = mouse_delta
        self.scroll_delta = scroll_delta
        self.mouse_left_held = mouse_left_held
        self.keys = set(keys)

    def to_c(self) -> CameraInput:
        mask = 0
        for k in self.keys:
            mask |= KEY_BITS[k.lower()]
        return CameraInput(float(self.mouse_delta[0]), float(self.mouse_delta[1]),
                           float(self.scroll_delta), int(bool(self.mouse_left_held)), mask)


class _Camera:
    """Camera (src/camera/camera.h:5-16) over ds_camera_*: update() each
    frame, then view_matrix / projection_matrix / position for the renderer."""
    _kind = -1
    _speed_names: dict = {}

    def __init__(self, aspect: float, fov_degrees: float = 40.0, near_plane: float = 0.1,
                 far_plane: float = 100.0):
        h = C.c_void_p()
        check(lib().ds_camera_create(self._kind, aspect, fov_degrees, near_plane, far_plane,
                                     C.byref(h)), "ds_camera_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().ds_camera_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def update(self, inp: Input, dt: float) -> bool:
        changed = C.c_int()
        c_in = inp.to_c()
        check(lib().ds_camera_update(self._h, C.byref(c_in), dt, C.byref(changed)),
              "ds_camera_update")
        return bool(changed.value)

    def set_aspect(self, aspect: float) -> None:
        check(lib().ds_camera_set_aspect(self._h, aspect), "ds_camera_set_aspect")

    def set_ortho_mode(self, enable: bool) -> None:
        check(lib().ds_camera_set_ortho(self._h, int(bool(enable))), "ds_camera_set_ortho")

    def matrices(self):
        """(view, proj, eye): column-major 4x4s as 16 floats, eye position."""
        v, p, e = np.empty(16, np.float32), np.empty(16, np.float32), np.empty(3, np.float32)
        check(lib().ds_camera_matrices(self._h, fptr(v), fptr(p), fptr(e)),
              "ds_camera_matrices")
        return v, p, e

    @property
    def view_matrix(self) -> np.ndarray:
        return self.matrices()[0]

    @property
    def projection_matrix(self) -> np.ndarray:
        return self.matrices()[1]

    @property
    def position(self) -> np.ndarray:
        return self.matrices()[2]

    def _speeds(self) -> CameraSpeeds:
        s = CameraSpeeds()
        check(lib().ds_camera_get_speeds(self._h, C.byref(s)), "ds_camera_get_speeds")
        return s

    def __getattr__(self, name):
        field = type(self)._speed_names.get(name)
        if field is None:
            raise AttributeError(name)
        return getattr(self._speeds(), field)

    def __setattr__(self, name, value):
        field = type(self)._speed_names.get(name)
        if field is None:
            object.__setattr__(self, name, value)
            return
        s = self._speeds()
        setattr(s, field, float(value))
        check(lib().ds_camera_set_speeds(self._h, C.byref(s)), "ds_camera_set_speeds")


class ArcballCamera(_Camera):
    """Orbit / shift-pan / scroll-zoom camera (src/camera/arcball_camera.h:20-71)."""
    _kind = 0
    _speed_names = {"speed_rotate": "rotate", "speed_pan": "pan", "speed_zoom": "zoom",
                    "zoom_min_dist": "zoom_min_dist", "zoom_max_dist": "zoom_max_dist"}


class FlyCamera(_Camera):
    """WASD/RF/QE + drag-look camera (src/camera/fly_camera.h:30-76)."""
    _kind = 1
    _speed_names = {"speed_move": "move", "speed_look": "look", "speed_roll": "roll",
                    "speed_mult_shift": "mult_shift", "speed_mult_ctrl": "mult_ctrl"}


def load_image(path: str, width: int, height: int, padding: int = 0):
    """ImageLoader::load: PNG / binary PPM fitted into a width x height canvas
    (aspect kept, centred, black elsewhere) -> (HWC floats in [0,1], src_w, src_h)."""
    out = np.empty(width * height * 3, np.float32)
    sw, sh = C.c_int32(), C.c_int32()
    check(lib().ds_image_load(str(path).encode(), width, height, padding, fptr(out),
                              C.byref(sw), C.byref(sh)), "ds_image_load")
    return out, sw.value, sh.value


def save_png(path: str, rgb: np.ndarray, width: int, height: int) -> None:
    """ImageSaver::saveAsPNG of HWC floats in [0,1]."""
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
    assert rgb.size == width * height * 3
    check(lib().ds_image_save_png(str(path).encode(), fptr(rgb), width, height),
          "ds_image_save_png")


def load_ply(path: str):
    """3DGS PLY -> (slab, count, sh_degree), the slab in this library's
    field-major layout and the OpenGL convention (see ds_ply_load)."""
    n, d = C.c_int64(), C.c_int32()
    p = str(path).encode()
    check(lib().ds_ply_info(p, C.byref(n), C.byref(d)), "ds_ply_info")
    slab = np.empty(slab_floats(n.value, d.value), np.float32)
    check(lib().ds_ply_load(p, fptr(slab), n.value, d.value), "ds_ply_load")
    return slab, n.value, d.value


def save_ply(path: str, slab: np.ndarray, count: int, sh_degree: int) -> None:
    slab = np.ascontiguousarray(slab, np.float32).reshape(-1)
    assert slab.size == slab_floats(count, sh_degree)
    check(lib().ds_ply_save(str(path).encode(), fptr(slab), count, sh_degree), "ds_ply_save")


def normalize_scene(slab: np.ndarray, count: int, scene_scale: float = 1.0) -> np.ndarray:
    """SplatUtils::normalizeScene on a host slab (in place; also returned)."""
    assert slab.dtype == np.float32 and slab.flags.c_contiguous
    check(lib().ds_normalize_scene(fptr(slab), count, scene_scale), "ds_normalize_scene")
    return slab


def psnr_from_loss(loss: float) -> float:
    """PSNR of the reference's loss (sum over channels / pixels): 10 log10(3 / loss)."""
    return 10.0 * math.log10(3.0 / max(loss, 1e-30))


# ---------------------------------------------------------------- objects
TILE_ORDERS = {0: "rows", 1: "heavy_first"}


def tile_order_name(raster_handle) -> str:
    """Compositing dispatch order of a ds_raster ("rows" or "heavy_first")."""
    o = C.c_int()
    check(lib().ds_raster_get_tile_order(raster_handle, C.byref(o)), "ds_raster_get_tile_order")
    return TILE_ORDERS[o.value]


DEPTH_SORTS = ("onesweep_8bit", "radix_8bit", "radix_11bit")


def depth_sort_name(raster_handle) -> str:
    """How a ds_raster sorts splats by depth (DS_DEPTH_SORT_*); the tile pass
    after it has 8-bit digits."""
    o = C.c_int()
    check(lib().ds_raster_get_depth_sort(raster_handle, C.byref(o)), "ds_raster_get_depth_sort")
    return DEPTH_SORTS[o.value]


def depth_sort_times(raster_handle) -> dict:
    """ms per sort of each depth-sort mode at the last autotune (None: not
    timed, or not available for the raster)."""
    ms = np.zeros(3, np.float32)
    check(lib().ds_raster_depth_sort_times(raster_handle, fptr(ms)), "ds_raster_depth_sort_times")
    return {n: (round(float(v), 4) if v >= 0 else None) for n, v in zip(DEPTH_SORTS, ms)}


class Rasterizer:
    """ds_raster: tile assign + sort + compositing, forward and backward."""

    def __init__(self, width: int, height: int, count_cap: int, isect_cap: int = 0,
                 keep_keys: bool = False, grad_mode: int = 0):
        cfg = RasterConfig(width, height, count_cap, isect_cap, int(keep_keys), int(grad_mode))
        h = C.c_void_p()
        check(lib().ds_raster_create(C.byref(cfg), C.byref(h)), "ds_raster_create")
        self.h = h.value
        self.width, self.height = width, height

    def forward(self, splat2d, out_rgb: DeviceArray, target=None, loss=None, grad=None,
                stream=None) -> None:
        check(lib().ds_raster_fwd(self.h, C.byref(splat2d), out_rgb.ptr,
                                  target.ptr if target else None, loss.ptr if loss else None,
                                  grad.ptr if grad else None, stream), "ds_raster_fwd")

    def backward(self, splat2d, grad_rgb: DeviceArray, grad_splat2d, stream=None) -> None:
        check(lib().ds_raster_bwd(self.h, C.byref(splat2d), grad_rgb.ptr,
                                  C.byref(grad_splat2d), stream), "ds_raster_bwd")

    def counts(self, stream=None):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().ds_raster_counts(self.h, stream, C.byref(a), C.byref(b), C.byref(c)),
              "ds_raster_counts")
        return a.value, b.value, c.value

    def state(self):
        """Downloads the forward state: dict of numpy arrays (synchronises)."""
        check(lib().ds_synchronize(), "ds_synchronize")
        v = RasterView()
        check(lib().ds_raster_get_view(self.h, C.byref(v)), "ds_raster_get_view")
        n_isects, visible, wanted = self.counts()
        ntiles = v.num_tiles_x * v.num_tiles_y
        npix = v.width * v.height
        st = {
            "n_isects": n_isects, "visible": visible, "wanted": wanted,
            "flatten_ids": download_ptr(v.flatten_ids, np.uint32, n_isects),
            "tile_ranges": download_ptr(v.tile_ranges, np.int32, 2 * ntiles),
            "T": download_ptr(v.transmittance, np.float32, npix),
            "last_ids": download_ptr(v.last_ids, np.int32, npix),
            "num_tiles_x": v.num_tiles_x, "num_tiles_y": v.num_tiles_y,
        }
        if v.tile_order:
            st["tile_order"] = download_ptr(v.tile_order, np.uint32, ntiles)
        if v.isect_keys_sorted:
            st["keys"] = download_ptr(v.isect_keys_sorted, np.uint64, n_isects)
        return st

    def close(self):
        if self.h:
            lib().ds_raster_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiRasterizer(Rasterizer):
    """ds_raster over several images at once: instance k owns the k-th run of
    `count` splats and its own width x height image; tile ids, pixel arrays
    and the sorted pair list are global (instances one after another)."""

    def __init__(self, sizes, isect_cap: int = 0, keep_keys: bool = False, grad_mode: int = 0):
        self.sizes = [(int(w), int(h), int(n)) for w, h, n in sizes]
        arr = (RasterInstance * len(self.sizes))(*[RasterInstance(w, h, n)
                                                    for w, h, n in self.sizes])
        hd = C.c_void_p()
        check(lib().ds_raster_create_multi(arr, len(self.sizes), isect_cap, int(keep_keys),
                                           int(grad_mode), C.byref(hd)), "ds_raster_create_multi")
        self.h = hd.value
        self.width = self.height = 0
        self.tile_base, self.pix_base, self.splat_base = [], [], []
        t = p = q = 0
        for w, h, n in self.sizes:
            self.tile_base.append(t)
            self.pix_base.append(p)
            self.splat_base.append(q)
            t += ((w + 15) // 16) * ((h + 15) // 16)
            p += w * h
            q += n
        self.num_tiles, self.num_pixels, self.num_splats = t, p, q

    def splat_instances(self) -> int:
        ptr = C.c_void_p()
        check(lib().ds_raster_splat_instances(self.h, C.byref(ptr)), "ds_raster_splat_instances")
        return ptr.value

    def state(self):
        """Forward state over all instances (synchronises)."""
        check(lib().ds_synchronize(), "ds_synchronize")
        v = RasterView()
        check(lib().ds_raster_get_view(self.h, C.byref(v)), "ds_raster_get_view")
        n_isects, visible, wanted = self.counts()
        st = {
            "n_isects": n_isects, "visible": visible, "wanted": wanted,
            "flatten_ids": download_ptr(v.flatten_ids, np.uint32, n_isects),
            "tile_ranges": download_ptr(v.tile_ranges, np.int32, 2 * self.num_tiles),
            "T": download_ptr(v.transmittance, np.float32, self.num_pixels),
            "last_ids": download_ptr(v.last_ids, np.int32, self.num_pixels),
        }
        if v.tile_order:
            st["tile_order"] = download_ptr(v.tile_order, np.uint32, self.num_tiles)
        if v.isect_keys_sorted:
            st["keys"] = download_ptr(v.isect_keys_sorted, np.uint64, n_isects)
        return st


class ImageFitter:
    """ds_fitter: the reference's ImageFitter (activ -> proj -> raster -> MSE -> Adam)."""

    def __init__(self, width: int, height: int, count: int, *, view=None, proj=None,
                 cam_pos=None, sh_degree: int = 0, lr=None, fused: bool = True,
                 isect_cap: int = 0, grad_mode: int = 0):
        cfg = FitterConfig()
        lib().ds_fitter_default_config(width, height, count, C.byref(cfg))
        cfg.sh_degree = sh_degree
        if view is not None:
            cfg.view[:] = [float(x) for x in np.asarray(view, np.float32).reshape(-1)]
        if proj is not None:
            cfg.proj[:] = [float(x) for x in np.asarray(proj, np.float32).reshape(-1)]
        if cam_pos is not None:
            cfg.cam_pos[:] = [float(x) for x in cam_pos]
        for k, v in (lr or {}).items():
            setattr(cfg, k, float(v))
        cfg.isect_cap = isect_cap
        cfg.fused = int(fused)
        cfg.grad_mode = int(grad_mode)
        h = C.c_void_p()
        check(lib().ds_fitter_create(C.byref(cfg), C.byref(h)), "ds_fitter_create")
        self.h = h.value
        self.cfg = cfg
        self.width, self.height, self.count, self.sh_degree = width, height, count, sh_degree

    def set_target(self, rgb: np.ndarray) -> None:
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
        assert rgb.size == self.width * self.height * 3
        check(lib().ds_fitter_set_target(self.h, fptr(rgb)), "ds_fitter_set_target")

    def upload_params(self, slab: np.ndarray) -> None:
        slab = np.ascontiguousarray(slab, np.float32).reshape(-1)
        assert slab.size == slab_floats(self.count, self.sh_degree)
        check(lib().ds_fitter_upload_params(self.h, fptr(slab)), "ds_fitter_upload_params")

    def params(self) -> np.ndarray:
        out = np.empty(slab_floats(self.count, self.sh_degree), np.float32)
        check(lib().ds_fitter_download_params(self.h, fptr(out)), "ds_fitter_download_params")
        return out

    def grads(self) -> np.ndarray:
        out = np.empty(slab_floats(self.count, self.sh_degree), np.float32)
        check(lib().ds_fitter_download_grads(self.h, fptr(out)), "ds_fitter_download_grads")
        return out

    def set_densify_stats(self, on: bool = True) -> None:
        """Accumulate per-splat gradient statistics in every later step."""
        check(lib().ds_fitter_set_densify_stats(self.h, int(on)), "ds_fitter_set_densify_stats")

    def densify_stats(self):
        """(sum of |dL/d(NDC centre)|, visits) per splat."""
        g = np.empty(self.count, np.float32)
        v = np.empty(self.count, np.uint32)
        check(lib().ds_fitter_densify_stats(self.h, fptr(g), v.ctypes.data),
              "ds_fitter_densify_stats")
        return g, v

    def densify(self, **cfg):
        """One clone / split / prune (ds_fitter_densify). Keyword arguments
        override ds_densify_default_config. Returns (result dict, per-splat
        action codes of the splats before)."""
        c = DensifyConfig()
        lib().ds_densify_default_config(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        r = DensifyResult()
        acts = np.empty(max(self.count, 1), np.uint8)
        check(lib().ds_fitter_densify(self.h, C.byref(c), C.byref(r), acts.ctypes.data),
              "ds_fitter_densify")
        n = C.c_int32()
        check(lib().ds_fitter_count(self.h, C.byref(n)), "ds_fitter_count")
        acts = acts[:self.count]
        self.count = n.value
        return {k: getattr(r, k) for k, _ in DensifyResult._fields_}, acts

    def moments(self):
        """Adam moments (m, v) in the parameter slab's layout."""
        m = np.empty(slab_floats(self.count, self.sh_degree), np.float32)
        v = np.empty_like(m)
        check(lib().ds_fitter_download_moments(self.h, fptr(m), fptr(v)),
              "ds_fitter_download_moments")
        return m, v

    def step_count(self) -> int:
        c = C.c_int32()
        check(lib().ds_fitter_step_count(self.h, C.byref(c)), "ds_fitter_step_count")
        return c.value

    def step(self, n: int = 1) -> None:
        check(lib().ds_fitter_step(self.h, n), "ds_fitter_step")

    def forward(self) -> None:
        check(lib().ds_fitter_forward(self.h), "ds_fitter_forward")

    def loss(self) -> float:
        v = C.c_float()
        check(lib().ds_fitter_loss(self.h, C.byref(v)), "ds_fitter_loss")
        return v.value

    def synchronize(self) -> None:
        check(lib().ds_fitter_synchronize(self.h), "ds_fitter_synchronize")

    def image(self) -> np.ndarray:
        out = np.empty(self.width * self.height * 3, np.float32)
        check(lib().ds_fitter_download_image(self.h, fptr(out)), "ds_fitter_download_image")
        return out

    def save_ply(self, path: str) -> None:
        """ImageFitter::savePLY."""
        check(lib().ds_fitter_save_ply(self.h, str(path).encode()), "ds_fitter_save_ply")

    def param_slabs(self):
        """Device pointers (parameter slab, parameter-gradient slab)."""
        p, g = C.c_void_p(), C.c_void_p()
        check(lib().ds_fitter_param_slabs(self.h, C.byref(p), C.byref(g)), "ds_fitter_param_slabs")
        return p.value, g.value

    def set_target_device(self, ptr) -> None:
        check(lib().ds_fitter_set_target_device(self.h, ptr), "ds_fitter_set_target_device")

    def reset_optimizer(self) -> None:
        check(lib().ds_fitter_reset_optimizer(self.h), "ds_fitter_reset_optimizer")

    def forward_backward(self) -> None:
        check(lib().ds_fitter_forward_backward(self.h), "ds_fitter_forward_backward")

    def set_graph(self, on: bool) -> None:
        check(lib().ds_fitter_set_graph(self.h, int(on)), "ds_fitter_set_graph")

    def prepare(self, n: int) -> None:
        """Graph mode: capture what step(n) replays, without running it."""
        check(lib().ds_fitter_prepare(self.h, n), "ds_fitter_prepare")

    def autotune(self, reps: int = 3) -> dict:
        """Picks the fastest compositing launch shapes on the current state."""
        f, b = C.c_int(), C.c_int()
        tf, tb = C.c_float(), C.c_float()
        check(lib().ds_fitter_autotune(self.h, reps, C.byref(f), C.byref(b), C.byref(tf),
                                       C.byref(tb)), "ds_fitter_autotune")
        r = C.c_void_p()
        check(lib().ds_fitter_get_raster(self.h, C.byref(r)), "ds_fitter_get_raster")
        return {"fwd_pix_per_thread": f.value, "bwd_pix_per_thread": b.value,
                "fwd_ms": round(tf.value, 4), "bwd_ms": round(tb.value, 4),
                "tile_order": tile_order_name(r), "depth_sort": depth_sort_name(r),
                "depth_sort_ms": depth_sort_times(r), "tile_radix_bits": 8}

    def set_tuning(self, fwd_ppt: int, bwd_ppt: int) -> None:
        r = C.c_void_p()
        check(lib().ds_fitter_get_raster(self.h, C.byref(r)), "ds_fitter_get_raster")
        check(lib().ds_raster_set_tuning(r, fwd_ppt, bwd_ppt), "ds_raster_set_tuning")

    def stream(self):
        s = C.c_void_p()
        check(lib().ds_fitter_get_stream(self.h, C.byref(s)), "ds_fitter_get_stream")
        return s.value

    def raster_state(self):
        r = C.c_void_p()
        check(lib().ds_fitter_get_raster(self.h, C.byref(r)), "ds_fitter_get_raster")
        ras = Rasterizer.__new__(Rasterizer)
        ras.h, ras.width, ras.height = r.value, self.width, self.height
        try:
            return ras.state()
        finally:
            ras.h = None

    def close(self):
        if self.h:
            lib().ds_fitter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class InitNet:
    """ds_initnet: the learned initializer, a bf16 MFMA CNN of six 3x3
    convolutions producing 16 SH-0 splats per 16x16-pixel cell (one per 4x4
    pixels), trained on (image, fitted splats) pairs or through the renderer."""

    HEAD_VALUES = 10  # per splat

    def __init__(self, width: int, height: int, fovy: float = math.radians(60.0), seed: int = 0):
        h = C.c_void_p()
        check(lib().ds_initnet_create(width, height, fovy, seed, C.byref(h)), "ds_initnet_create")
        self.h = h.value
        self.width, self.height = width, height
        npar, nspl = C.c_int64(), C.c_int32()
        check(lib().ds_initnet_info(self.h, C.byref(npar), C.byref(nspl)), "ds_initnet_info")
        self.num_params, self.num_splats = npar.value, nspl.value

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(lib().ds_initnet_get_params(self.h, fptr(out)), "ds_initnet_get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        assert p.size == self.num_params
        check(lib().ds_initnet_set_params(self.h, fptr(p)), "ds_initnet_set_params")

    def infer(self, image, slab, stream=None) -> None:
        """image, slab: DeviceArray or device pointers."""
        ip = getattr(image, "ptr", image)
        sp = getattr(slab, "ptr", slab)
        check(lib().ds_initnet_infer(self.h, ip, sp, stream), "ds_initnet_infer")

    def head(self) -> np.ndarray:
        """Head values of the last forward: [cells][16 splats x 10]."""
        out = np.empty(self.num_splats * self.HEAD_VALUES, np.float32)
        check(lib().ds_initnet_head(self.h, fptr(out)), "ds_initnet_head")
        return out

    def head_targets(self, slab, targets, stream=None) -> None:
        """Regression targets (device, head layout) of a fitted slab (device)."""
        check(lib().ds_initnet_head_targets(self.h, getattr(slab, "ptr", slab),
                                            getattr(targets, "ptr", targets), stream),
              "ds_initnet_head_targets")

    def set_loss_weights(self, weights=None) -> None:
        """Per head value weights of the supervised loss (None: defaults)."""
        if weights is None:
            check(lib().ds_initnet_set_loss_weights(self.h, None), "ds_initnet_set_loss_weights")
            return
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        assert w.size == self.HEAD_VALUES
        check(lib().ds_initnet_set_loss_weights(self.h, fptr(w)), "ds_initnet_set_loss_weights")

    def train_step_supervised(self, image, targets, lr: float, loss=None, stream=None) -> None:
        check(lib().ds_initnet_train_step_supervised(
            self.h, getattr(image, "ptr", image), getattr(targets, "ptr", targets), lr,
            getattr(loss, "ptr", loss), stream), "ds_initnet_train_step_supervised")

    def grads(self, image: "DeviceArray", grad_slab: "DeviceArray") -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(lib().ds_initnet_grads(self.h, image.ptr, grad_slab.ptr, fptr(out)),
              "ds_initnet_grads")
        return out

    def train_step(self, fitter: "ImageFitter", image: "DeviceArray", lr: float) -> None:
        check(lib().ds_initnet_train_step(self.h, fitter.h, image.ptr, lr),
              "ds_initnet_train_step")

    def close(self):
        if self.h:
            lib().ds_initnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchFitter:
    """ds_batch: many ImageFitters stepped together, each replaying its own
    captured step, spread over a few streams. Keeps the fitters alive."""

    def __init__(self, fitters):
        self.fitters = list(fitters)
        arr = (C.c_void_p * len(self.fitters))(*[f.h for f in self.fitters])
        h = C.c_void_p()
        check(lib().ds_batch_create(arr, len(self.fitters), C.byref(h)), "ds_batch_create")
        self.h = h.value

    def step(self, n: int = 1) -> None:
        check(lib().ds_batch_step(self.h, n), "ds_batch_step")

    def synchronize(self) -> None:
        check(lib().ds_batch_synchronize(self.h), "ds_batch_synchronize")

    def set_streams(self, n: int) -> None:
        check(lib().ds_batch_set_streams(self.h, n), "ds_batch_set_streams")

    def autotune(self, reps: int = 3) -> dict:
        ns = C.c_int()
        ms = C.c_float()
        check(lib().ds_batch_autotune(self.h, reps, C.byref(ns), C.byref(ms)), "ds_batch_autotune")
        return {"streams": ns.value, "ms_per_batch_step": round(ms.value, 4)}

    def close(self):
        if self.h:
            lib().ds_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiFitter:
    """ds_multifit: many fits as one, every phase of a step one launch over
    all instances (and the whole step one HIP graph when graphs are on).
    instances: (width, height, count, view, proj, cam_pos) per fit."""

    def __init__(self, instances, *, sh_degree: int = 0, lr=None, isect_cap: int = 0,
                 grad_mode: int = 0):
        self.inst = [(int(w), int(h), int(n)) for w, h, n, *_ in instances]
        arr = (MultiFitInstance * len(instances))()
        for k, (w, h, n, view, proj, cam) in enumerate(instances):
            q = arr[k]
            q.width, q.height, q.count = int(w), int(h), int(n)
            q.view[:] = [float(x) for x in np.asarray(view, np.float32).reshape(-1)]
            q.proj[:] = [float(x) for x in np.asarray(proj, np.float32).reshape(-1)]
            q.cam_pos[:] = [float(x) for x in cam]
        cfg = FitterConfig()
        lib().ds_fitter_default_config(1, 1, 0, C.byref(cfg))
        cfg.sh_degree = sh_degree
        cfg.grad_mode = int(grad_mode)
        for k, v in (lr or {}).items():
            setattr(cfg, k, float(v))
        hd = C.c_void_p()
        check(lib().ds_multifit_create(arr, len(instances), C.byref(cfg), isect_cap,
                                       C.byref(hd)), "ds_multifit_create")
        self.h = hd.value
        self.sh_degree = sh_degree

    def __len__(self) -> int:
        return len(self.inst)

    def set_target(self, k: int, rgb: np.ndarray) -> None:
        w, h, _ = self.inst[k]
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
        assert rgb.size == 3 * w * h
        check(lib().ds_multifit_set_target(self.h, k, fptr(rgb)), "ds_multifit_set_target")

    def upload_params(self, k: int, slab: np.ndarray) -> None:
        slab = np.ascontiguousarray(slab, np.float32).reshape(-1)
        assert slab.size == slab_floats(self.inst[k][2], self.sh_degree)
        check(lib().ds_multifit_upload_params(self.h, k, fptr(slab)), "ds_multifit_upload_params")

    def params(self, k: int) -> np.ndarray:
        out = np.empty(slab_floats(self.inst[k][2], self.sh_degree), np.float32)
        check(lib().ds_multifit_download_params(self.h, k, fptr(out)),
              "ds_multifit_download_params")
        return out

    def grads(self, k: int) -> np.ndarray:
        out = np.empty(slab_floats(self.inst[k][2], self.sh_degree), np.float32)
        check(lib().ds_multifit_download_grads(self.h, k, fptr(out)), "ds_multifit_download_grads")
        return out

    def image(self, k: int) -> np.ndarray:
        w, h, _ = self.inst[k]
        out = np.empty(3 * w * h, np.float32)
        check(lib().ds_multifit_download_image(self.h, k, fptr(out)), "ds_multifit_download_image")
        return out

    def step(self, n: int = 1) -> None:
        check(lib().ds_multifit_step(self.h, n), "ds_multifit_step")

    def forward(self) -> None:
        check(lib().ds_multifit_forward(self.h), "ds_multifit_forward")

    def forward_backward(self) -> None:
        check(lib().ds_multifit_forward_backward(self.h), "ds_multifit_forward_backward")

    def losses(self) -> np.ndarray:
        out = np.empty(len(self.inst), np.float32)
        check(lib().ds_multifit_losses(self.h, fptr(out)), "ds_multifit_losses")
        return out

    def synchronize(self) -> None:
        check(lib().ds_multifit_synchronize(self.h), "ds_multifit_synchronize")

    def set_graph(self, on: bool) -> None:
        check(lib().ds_multifit_set_graph(self.h, int(on)), "ds_multifit_set_graph")

    def autotune(self, reps: int = 3) -> dict:
        f, b = C.c_int(), C.c_int()
        tf, tb = C.c_float(), C.c_float()
        check(lib().ds_multifit_autotune(self.h, reps, C.byref(f), C.byref(b), C.byref(tf),
                                         C.byref(tb)), "ds_multifit_autotune")
        r = C.c_void_p()
        check(lib().ds_multifit_get_raster(self.h, C.byref(r)), "ds_multifit_get_raster")
        return {"fwd_pix_per_thread": f.value, "bwd_pix_per_thread": b.value,
                "fwd_ms": round(tf.value, 4), "bwd_ms": round(tb.value, 4),
                "tile_order": tile_order_name(r), "depth_sort": depth_sort_name(r),
                "depth_sort_ms": depth_sort_times(r), "tile_radix_bits": 8}

    def set_tuning(self, fwd_ppt: int, bwd_ppt: int) -> None:
        r = C.c_void_p()
        check(lib().ds_multifit_get_raster(self.h, C.byref(r)), "ds_multifit_get_raster")
        check(lib().ds_raster_set_tuning(r, fwd_ppt, bwd_ppt), "ds_raster_set_tuning")

    def step_count(self) -> int:
        s = C.c_int32()
        check(lib().ds_multifit_step_count(self.h, C.byref(s)), "ds_multifit_step_count")
        return s.value

    def stream(self):
        s = C.c_void_p()
        check(lib().ds_multifit_get_stream(self.h, C.byref(s)), "ds_multifit_get_stream")
        return s.value

    def bases(self):
        """Per instance: its first global tile, first pixel and first splat."""
        tb, pb, sb = [], [], []
        t = p = q = 0
        for w, h, n in self.inst:
            tb.append(t)
            pb.append(p)
            sb.append(q)
            t += ((w + 15) // 16) * ((h + 15) // 16)
            p += w * h
            q += n
        return tb, pb, sb

    def raster_state(self):
        """Forward state of the last step over all instances (synchronises)."""
        r = C.c_void_p()
        check(lib().ds_multifit_get_raster(self.h, C.byref(r)), "ds_multifit_get_raster")
        ras = MultiRasterizer.__new__(MultiRasterizer)
        ras.sizes = list(self.inst)
        ras.tile_base, ras.pix_base, ras.splat_base = self.bases()
        ras.num_tiles = sum(((w + 15) // 16) * ((h + 15) // 16) for w, h, _ in self.inst)
        ras.num_pixels = sum(w * h for w, h, _ in self.inst)
        ras.num_splats = sum(n for _, _, n in self.inst)
        ras.h, ras.width, ras.height = r.value, 0, 0
        try:
            return ras.state()
        finally:
            ras.h = None

    def views(self):
        """Device views (Splat2D, dL/dSplat2D, dL/d(image) pointer) over all
        instances."""
        g, s2, gs2, grad = Gauss3D(), Splat2D(), Splat2D(), C.c_void_p()
        check(lib().ds_multifit_views(self.h, C.byref(g), C.byref(s2), C.byref(gs2),
                                      C.byref(grad)), "ds_multifit_views")
        return s2, gs2, grad.value

    def raster_counts(self):
        r = C.c_void_p()
        check(lib().ds_multifit_get_raster(self.h, C.byref(r)), "ds_multifit_get_raster")
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().ds_raster_counts(r, self.stream(), C.byref(a), C.byref(b), C.byref(c)),
              "ds_raster_counts")
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            lib().ds_multifit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_instances(count: int = 64, seed: int = 3000, min_area: int = 128 * 128,
                    max_area: int = 1024 * 576, min_splats: int = 2000, max_splats: int = 100_000):
    """Sizes of the batched-fitting workload: per instance an image area
    log-uniform in [min_area, max_area] with an aspect ratio uniform in
    [1, 16/9], and a splat count log-uniform in [min_splats, max_splats].
    Returns (width, height, splats, seed) per instance."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        area = math.exp(rng.uniform(math.log(min_area), math.log(max_area)))
        aspect = rng.uniform(1.0, 16.0 / 9.0)
        w = max(16, int(round(math.sqrt(area * aspect))))
        h = max(16, int(round(area / w)))
        n = int(round(math.exp(rng.uniform(math.log(min_splats), math.log(max_splats)))))
        out.append((w, h, n, seed + i))
    return out


class SplatRenderer:
    """ds_renderer: the reference's SplatRenderer (forward only, SH degree 0-3)."""

    def __init__(self, width: int, height: int, count: int, sh_degree: int,
                 isect_cap: int = 0):
        h = C.c_void_p()
        check(lib().ds_renderer_create(width, height, count, sh_degree, isect_cap, C.byref(h)),
              "ds_renderer_create")
        self.h = h.value
        self.width, self.height, self.count, self.sh_degree = width, height, count, sh_degree

    @classmethod
    def from_ply(cls, path: str, width: int, height: int, scene_scale: float = 1.0,
                 isect_cap: int = 0) -> "SplatRenderer":
        """SplatRenderer::loadPLY: the scene of a 3DGS PLY, normalised."""
        h, n, d = C.c_void_p(), C.c_int64(), C.c_int32()
        check(lib().ds_renderer_load_ply(str(path).encode(), width, height, scene_scale,
                                         isect_cap, C.byref(h), C.byref(n), C.byref(d)),
              "ds_renderer_load_ply")
        r = cls.__new__(cls)
        r.h = h.value
        r.width, r.height, r.count, r.sh_degree = width, height, n.value, d.value
        return r

    def upload_params(self, slab: np.ndarray) -> None:
        slab = np.ascontiguousarray(slab, np.float32).reshape(-1)
        assert slab.size == slab_floats(self.count, self.sh_degree)
        check(lib().ds_renderer_upload_params(self.h, fptr(slab)), "ds_renderer_upload_params")

    def render(self, view, proj, cam_pos) -> None:
        v = np.ascontiguousarray(view, np.float32)
        p = np.ascontiguousarray(proj, np.float32)
        c = np.ascontiguousarray(cam_pos, np.float32)
        check(lib().ds_renderer_render(self.h, fptr(v), fptr(p), fptr(c)), "ds_renderer_render")

    def autotune(self, reps: int = 5) -> dict:
        """Times the last camera's frame for each compositing launch shape and
        keeps the fastest (needs one render first)."""
        best = C.c_int()
        ms = C.c_float()
        check(lib().ds_renderer_autotune(self.h, reps, C.byref(best), C.byref(ms)),
              "ds_renderer_autotune")
        r = C.c_void_p()
        check(lib().ds_renderer_get_raster(self.h, C.byref(r)), "ds_renderer_get_raster")
        return {"fwd_pix_per_thread": best.value, "ms_per_frame": round(ms.value, 4),
                "tile_order": tile_order_name(r), "depth_sort": depth_sort_name(r),
                "depth_sort_ms": depth_sort_times(r), "tile_radix_bits": 8}

    def synchronize(self) -> None:
        check(lib().ds_renderer_synchronize(self.h), "ds_renderer_synchronize")

    def image(self) -> np.ndarray:
        out = np.empty(self.width * self.height * 3, np.float32)
        check(lib().ds_renderer_download_image(self.h, fptr(out)), "ds_renderer_download_image")
        return out

    def stream(self):
        s = C.c_void_p()
        check(lib().ds_renderer_get_stream(self.h, C.byref(s)), "ds_renderer_get_stream")
        return s.value

    def raster_state(self):
        r = C.c_void_p()
        check(lib().ds_renderer_get_raster(self.h, C.byref(r)), "ds_renderer_get_raster")
        ras = Rasterizer.__new__(Rasterizer)
        ras.h, ras.width, ras.height = r.value, self.width, self.height
        try:
            return ras.state()
        finally:
            ras.h = None

    def close(self):
        if self.h:
            lib().ds_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
