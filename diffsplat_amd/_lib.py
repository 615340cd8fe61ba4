"""ctypes binding of libdiffsplat.so (the C-ABI declared in include/ds_api.h).

The engine's compute path is the HIP library; this module only declares its
entry points and moves numpy arrays in and out of device memory. It raises
ImportError if the library is not built -- there is no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffsplat.so")

DS_OK = 0
STATUS_NAMES = {0: "DS_OK", 1: "DS_ERR_INVALID_ARG", 2: "DS_ERR_HIP",
                3: "DS_ERR_OUT_OF_MEMORY", 4: "DS_ERR_CAPACITY", 5: "DS_ERR_STATE"}

GAUSS_FIELDS = ("pos_x", "pos_y", "pos_z", "scale_x", "scale_y", "scale_z",
                "rot_w", "rot_x", "rot_y", "rot_z", "sh_dc_r", "sh_dc_g", "sh_dc_b",
                "sh_rest_r", "sh_rest_g", "sh_rest_b", "logit_opacity")
SPLAT3D_FIELDS = ("pos_x", "pos_y", "pos_z", "cov_xx", "cov_xy", "cov_xz",
                  "cov_yy", "cov_yz", "cov_zz", "color_r", "color_g", "color_b", "opacity")
SPLAT2D_FIELDS = ("pos_x", "pos_y", "pos_z", "cov_xx", "cov_xy", "cov_yy",
                  "color_r", "color_g", "color_b", "opacity")


class DiffsplatError(RuntimeError):
    """A C-ABI call returned a status other than DS_OK."""

    def __init__(self, status: int, what: str, detail: str):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}: {detail}")


class Gauss3D(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in GAUSS_FIELDS] + [("count", C.c_int32),
                                                          ("sh_num_bands", C.c_int32)]


class Splat3D(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in SPLAT3D_FIELDS] + [("count", C.c_int32)]


class Splat2D(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in SPLAT2D_FIELDS] + [("count", C.c_int32)]


class RasterConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("count_cap", C.c_int32),
                ("isect_cap", C.c_int64), ("keep_keys", C.c_int32), ("grad_mode", C.c_int32)]


class RasterView(C.Structure):
    _fields_ = [("isect_keys_sorted", C.c_void_p), ("flatten_ids", C.c_void_p),
                ("tile_ranges", C.c_void_p), ("transmittance", C.c_void_p),
                ("last_ids", C.c_void_p), ("n_isects", C.c_void_p),
                ("n_isects_wanted", C.c_void_p), ("visible_count", C.c_void_p),
                ("num_tiles_x", C.c_int32), ("num_tiles_y", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("tile_order", C.c_void_p),
                ("lookback_error", C.c_void_p)]


class AdamGroup(C.Structure):
    _fields_ = [("param", C.c_void_p), ("grad", C.c_void_p), ("m", C.c_void_p),
                ("v", C.c_void_p), ("n", C.c_int64), ("lr", C.c_float),
                ("decay", C.c_float)]


class DensifyConfig(C.Structure):
    _fields_ = [("grad_threshold", C.c_float), ("percent_dense", C.c_float),
                ("scene_extent", C.c_float), ("min_opacity", C.c_float),
                ("max_scale_frac", C.c_float), ("split_scale_div", C.c_float),
                ("seed", C.c_uint64)]


class DensifyResult(C.Structure):
    _fields_ = [("count_before", C.c_int32), ("count_after", C.c_int32), ("pruned", C.c_int32),
                ("cloned", C.c_int32), ("split", C.c_int32)]


class AdamHParams(C.Structure):
    _fields_ = [("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float)]


class FitterConfig(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("count", C.c_int32),
                ("sh_degree", C.c_int32), ("view", C.c_float * 16), ("proj", C.c_float * 16),
                ("cam_pos", C.c_float * 3), ("lr_pos", C.c_float), ("lr_scale", C.c_float),
                ("lr_rot", C.c_float), ("lr_sh_dc", C.c_float), ("lr_sh_rest", C.c_float),
                ("lr_opacity", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("isect_cap", C.c_int64), ("fused", C.c_int32),
                ("grad_mode", C.c_int32), ("lr_pos_decay", C.c_float)]


class RasterInstance(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("count", C.c_int32)]


class MultiFitInstance(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("count", C.c_int32),
                ("view", C.c_float * 16), ("proj", C.c_float * 16), ("cam_pos", C.c_float * 3)]


class CameraInput(C.Structure):
    _fields_ = [("mouse_dx", C.c_float), ("mouse_dy", C.c_float), ("scroll", C.c_float),
                ("mouse_left", C.c_int32), ("keys", C.c_uint32)]


class CameraSpeeds(C.Structure):
    _fields_ = [(f, C.c_float) for f in ("rotate", "pan", "zoom", "zoom_min_dist",
                                         "zoom_max_dist", "move", "look", "roll",
                                         "mult_shift", "mult_ctrl")]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_SIGS = {
    "ds_status_string": (C.c_char_p, [C.c_int]),
    "ds_last_error": (C.c_char_p, []),
    "ds_version": (C.c_char_p, []),
    "ds_device_count": (C.c_int, []),
    "ds_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "ds_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "ds_free": (C.c_int, [_P]),
    "ds_memcpy_htod": (C.c_int, [_P, _P, C.c_size_t]),
    "ds_memcpy_dtoh": (C.c_int, [_P, _P, C.c_size_t]),
    "ds_memcpy_dtod": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "ds_fill_u32": (C.c_int, [_P, C.c_uint32, C.c_size_t, _P]),
    "ds_synchronize": (C.c_int, []),
    "ds_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "ds_stream_destroy": (C.c_int, [_P]),
    "ds_stream_synchronize": (C.c_int, [_P]),
    "ds_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "ds_event_destroy": (C.c_int, [_P]),
    "ds_event_record": (C.c_int, [_P, _P]),
    "ds_event_synchronize": (C.c_int, [_P]),
    "ds_event_elapsed_ms": (C.c_int, [_P, _P, _F]),
    "ds_activ_fwd": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat3D), C.c_int, _F, _P]),
    "ds_activ_bwd": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat3D), C.POINTER(Splat3D),
                               C.POINTER(Gauss3D), C.c_int, _F, C.c_int, _P]),
    "ds_project_fwd": (C.c_int, [C.POINTER(Splat3D), C.POINTER(Splat2D), _F, _P]),
    "ds_project_bwd": (C.c_int, [C.POINTER(Splat3D), C.POINTER(Splat2D), C.POINTER(Splat3D),
                                 _F, C.c_int, _P]),
    "ds_activ_project_fwd": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat2D), C.c_int, _F,
                                       _F, _P]),
    "ds_activ_project_bwd": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat2D),
                                       C.POINTER(Gauss3D), C.c_int, _F, _F, C.c_int, _P]),
    "ds_activ_project_fwd_multi": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat2D), C.c_int,
                                             _P, _P, _P]),
    "ds_activ_project_bwd_multi": (C.c_int, [C.POINTER(Gauss3D), C.POINTER(Splat2D),
                                             C.POINTER(Gauss3D), C.c_int, _P, _P, C.c_int,
                                             _P]),
    "ds_raster_create": (C.c_int, [C.POINTER(RasterConfig), C.POINTER(C.c_void_p)]),
    "ds_raster_create_multi": (C.c_int, [C.POINTER(RasterInstance), C.c_int, C.c_int64,
                                         C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "ds_raster_splat_instances": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_raster_destroy": (C.c_int, [_P]),
    "ds_raster_resize": (C.c_int, [_P, C.c_int, C.c_int]),
    "ds_raster_fwd": (C.c_int, [_P, C.POINTER(Splat2D), _P, _P, _P, _P, _P]),
    "ds_raster_bwd": (C.c_int, [_P, C.POINTER(Splat2D), _P, C.POINTER(Splat2D), _P]),
    "ds_raster_fwd_bwd": (C.c_int, [_P, C.POINTER(Splat2D), _P, _P, _P, _P, C.POINTER(Splat2D),
                                    _P]),
    "ds_raster_composite_fwd": (C.c_int, [_P, C.POINTER(Splat2D), _P, _P]),
    "ds_raster_release_backward": (C.c_int, [_P]),
    "ds_raster_set_depth_sort": (C.c_int, [_P, C.c_int]),
    "ds_raster_get_depth_sort": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ds_raster_sort": (C.c_int, [_P, C.POINTER(Splat2D), _P]),
    "ds_raster_autotune_depth_sort": (C.c_int, [_P, C.POINTER(Splat2D), C.c_int, _P,
                                                C.POINTER(C.c_int), _P]),
    "ds_raster_depth_sort_times": (C.c_int, [_P, _P]),
    "ds_raster_composite_fwd_bwd": (C.c_int, [_P, C.POINTER(Splat2D), _P, _P, _P, _P,
                                              C.POINTER(Splat2D), _P]),
    "ds_splat2d_zero": (C.c_int, [C.POINTER(Splat2D), _P]),
    "ds_raster_counts": (C.c_int, [_P, _P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]),
    "ds_raster_get_view": (C.c_int, [_P, C.POINTER(RasterView)]),
    "ds_raster_set_tuning": (C.c_int, [_P, C.c_int, C.c_int]),
    "ds_raster_get_tuning": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ds_raster_set_tile_order": (C.c_int, [_P, C.c_int]),
    "ds_raster_get_tile_order": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ds_fitter_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _F,
                                     _F]),
    "ds_sort_pairs_u64": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int, _P]),
    "ds_sort_pairs_u64_noalloc": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int, _P]),
    "ds_mse_fwd": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "ds_mse_bwd": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "ds_adam_step": (C.c_int, [C.POINTER(AdamGroup), C.c_int, C.POINTER(AdamHParams),
                               C.c_int32, _P]),
    "ds_adam_step_dev": (C.c_int, [C.POINTER(AdamGroup), C.c_int, C.POINTER(AdamHParams),
                                   _P, _P]),
    "ds_fitter_default_config": (None, [C.c_int, C.c_int, C.c_int, C.POINTER(FitterConfig)]),
    "ds_fitter_create": (C.c_int, [C.POINTER(FitterConfig), C.POINTER(C.c_void_p)]),
    "ds_fitter_destroy": (C.c_int, [_P]),
    "ds_fitter_set_target": (C.c_int, [_P, _F]),
    "ds_fitter_upload_params": (C.c_int, [_P, _F]),
    "ds_fitter_download_params": (C.c_int, [_P, _F]),
    "ds_fitter_download_grads": (C.c_int, [_P, _F]),
    "ds_fitter_download_moments": (C.c_int, [_P, _F, _F]),
    "ds_densify_default_config": (None, [C.POINTER(DensifyConfig)]),
    "ds_fitter_set_densify_stats": (C.c_int, [_P, C.c_int]),
    "ds_fitter_densify_stats": (C.c_int, [_P, _F, _P]),
    "ds_fitter_densify": (C.c_int, [_P, C.POINTER(DensifyConfig), C.POINTER(DensifyResult),
                                    _P]),
    "ds_fitter_count": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "ds_fitter_step": (C.c_int, [_P, C.c_int]),
    "ds_fitter_forward": (C.c_int, [_P]),
    "ds_fitter_loss": (C.c_int, [_P, _F]),
    "ds_fitter_synchronize": (C.c_int, [_P]),
    "ds_fitter_download_image": (C.c_int, [_P, _F]),
    "ds_fitter_get_raster": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_fitter_get_stream": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_fitter_set_graph": (C.c_int, [_P, C.c_int]),
    "ds_fitter_prepare": (C.c_int, [_P, C.c_int]),
    "ds_fitter_step_count": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "ds_fitter_views": (C.c_int, [_P, C.POINTER(Gauss3D), C.POINTER(Splat2D), C.POINTER(Splat2D),
                                  C.POINTER(C.c_void_p)]),
    "ds_renderer_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                                     C.POINTER(C.c_void_p)]),
    "ds_renderer_destroy": (C.c_int, [_P]),
    "ds_renderer_upload_params": (C.c_int, [_P, _F]),
    "ds_image_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ds_image_load": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, _F,
                                C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ds_image_save_png": (C.c_int, [C.c_char_p, _F, C.c_int, C.c_int]),
    "ds_image_save_png_device": (C.c_int, [C.c_char_p, _P, C.c_int, C.c_int]),
    "ds_ply_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "ds_ply_load": (C.c_int, [C.c_char_p, _F, C.c_int64, C.c_int32]),
    "ds_ply_save": (C.c_int, [C.c_char_p, _F, C.c_int64, C.c_int32]),
    "ds_normalize_scene": (C.c_int, [_F, C.c_int64, C.c_float]),
    "ds_renderer_load_ply": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_float, C.c_int64,
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int32)]),
    "ds_fitter_save_ply": (C.c_int, [_P, C.c_char_p]),
    "ds_initnet_create": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_void_p)]),
    "ds_initnet_destroy": (C.c_int, [_P]),
    "ds_initnet_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "ds_initnet_get_params": (C.c_int, [_P, _F]),
    "ds_initnet_set_params": (C.c_int, [_P, _F]),
    "ds_initnet_infer": (C.c_int, [_P, _P, _P, _P]),
    "ds_initnet_head": (C.c_int, [_P, _F]),
    "ds_initnet_grads": (C.c_int, [_P, _P, _P, _F]),
    "ds_initnet_train_step": (C.c_int, [_P, _P, _P, C.c_float]),
    "ds_initnet_head_targets": (C.c_int, [_P, _P, _P, _P]),
    "ds_initnet_train_step_supervised": (C.c_int, [_P, _P, _P, C.c_float, _P, _P]),
    "ds_initnet_set_loss_weights": (C.c_int, [_P, _P]),
    "ds_fitter_forward_backward": (C.c_int, [_P]),
    "ds_fitter_param_slabs": (C.c_int, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "ds_fitter_set_target_device": (C.c_int, [_P, _P]),
    "ds_fitter_reset_optimizer": (C.c_int, [_P]),
    "ds_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "ds_batch_destroy": (C.c_int, [_P]),
    "ds_batch_step": (C.c_int, [_P, C.c_int]),
    "ds_batch_synchronize": (C.c_int, [_P]),
    "ds_batch_set_streams": (C.c_int, [_P, C.c_int]),
    "ds_batch_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), _F]),
    "ds_multifit_create": (C.c_int, [C.POINTER(MultiFitInstance), C.c_int,
                                     C.POINTER(FitterConfig), C.c_int64,
                                     C.POINTER(C.c_void_p)]),
    "ds_multifit_destroy": (C.c_int, [_P]),
    "ds_multifit_set_target": (C.c_int, [_P, C.c_int, _F]),
    "ds_multifit_upload_params": (C.c_int, [_P, C.c_int, _F]),
    "ds_multifit_download_params": (C.c_int, [_P, C.c_int, _F]),
    "ds_multifit_download_grads": (C.c_int, [_P, C.c_int, _F]),
    "ds_multifit_download_image": (C.c_int, [_P, C.c_int, _F]),
    "ds_multifit_step": (C.c_int, [_P, C.c_int]),
    "ds_multifit_forward": (C.c_int, [_P]),
    "ds_multifit_forward_backward": (C.c_int, [_P]),
    "ds_multifit_losses": (C.c_int, [_P, _F]),
    "ds_multifit_synchronize": (C.c_int, [_P]),
    "ds_multifit_set_graph": (C.c_int, [_P, C.c_int]),
    "ds_multifit_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       _F, _F]),
    "ds_multifit_step_count": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "ds_multifit_get_raster": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_multifit_get_stream": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_multifit_views": (C.c_int, [_P, C.POINTER(Gauss3D), C.POINTER(Splat2D),
                                    C.POINTER(Splat2D), C.POINTER(C.c_void_p)]),
    "ds_renderer_render": (C.c_int, [_P, _F, _F, _F]),
    "ds_renderer_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), _F]),
    "ds_renderer_download_image": (C.c_int, [_P, _F]),
    "ds_renderer_views": (C.c_int, [_P, C.POINTER(Splat2D)]),
    "ds_renderer_get_raster": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_renderer_get_stream": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "ds_renderer_synchronize": (C.c_int, [_P]),
    "ds_gen_target_image": (None, [C.c_uint64, C.c_int, C.c_int, C.c_int, _F]),
    "ds_gen_frustum_splats": (None, [C.c_uint64, C.c_int, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_float, C.c_float, _F]),
    "ds_gen_cluster_scene": (None, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_float,
                                    C.c_float, _F]),
    "ds_gen_fitter_init": (None, [C.c_uint64, C.c_int, C.c_int, C.c_int, _F]),
    "ds_mat_perspective": (None, [C.c_float, C.c_float, C.c_float, C.c_float, _F]),
    "ds_mat_ortho": (None, [C.c_float] * 6 + [_F]),
    "ds_mat_look_at": (None, [_F, _F, _F, _F]),
    "ds_mat_mul": (None, [_F, _F, _F]),
    "ds_camera_default_speeds": (None, [C.POINTER(CameraSpeeds)]),
    "ds_camera_create": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.POINTER(_P)]),
    "ds_camera_destroy": (None, [_P]),
    "ds_camera_update": (C.c_int, [_P, C.POINTER(CameraInput), C.c_float,
                                   C.POINTER(C.c_int)]),
    "ds_camera_set_aspect": (C.c_int, [_P, C.c_float]),
    "ds_camera_set_ortho": (C.c_int, [_P, C.c_int]),
    "ds_camera_get_speeds": (C.c_int, [_P, C.POINTER(CameraSpeeds)]),
    "ds_camera_set_speeds": (C.c_int, [_P, C.POINTER(CameraSpeeds)]),
    "ds_camera_matrices": (C.c_int, [_P, _F, _F, _F]),
    "ds_slab_floats": (C.c_int64, [C.c_int, C.c_int]),
    "ds_sh_bands": (C.c_int, [C.c_int]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libdiffsplat.so once; raises ImportError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `make -C diffsplat_amd/csrc`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def declared_symbols() -> tuple:
    return tuple(_SIGS)


def check(status: int, what: str) -> None:
    if status != DS_OK:
        raise DiffsplatError(status, what, lib().ds_last_error().decode())


def fptr(a: np.ndarray):
    """float* of a C-contiguous float32 numpy array."""
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_F)


def farr(vals) -> C.Array:
    vals = list(vals)
    return (C.c_float * len(vals))(*vals)


def has_gpu() -> bool:
    try:
        return lib().ds_device_count() > 0
    except ImportError:
        return False


class DeviceArray:
    """A device allocation made by the engine's own allocator."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().ds_malloc(C.byref(p), max(self.nbytes, 4)), "ds_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().ds_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "ds_memcpy_htod")

    def download(self, dtype, count: int, offset_bytes: int = 0) -> np.ndarray:
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes:
            check(lib().ds_memcpy_dtoh(out.ctypes.data, self.ptr + offset_bytes, out.nbytes),
                  "ds_memcpy_dtoh")
        return out

    def zero(self) -> None:
        check(lib().ds_fill_u32(self.ptr, 0, self.nbytes // 4, None), "ds_fill_u32")
        check(lib().ds_synchronize(), "ds_synchronize")

    def free(self) -> None:
        if self.ptr:
            lib().ds_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ptr: int, dtype, count: int) -> np.ndarray:
    out = np.empty(int(count), dtype=dtype)
    if out.nbytes:
        check(lib().ds_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes), "ds_memcpy_dtoh")
    return out


def gauss_view(slab: DeviceArray, n: int, sh_degree: int) -> Gauss3D:
    """Gauss3D view over a field-major parameter slab (layout of ds_fitter_upload_params)."""
    nb = lib().ds_sh_bands(sh_degree)
    base = slab.ptr
    ptrs = [base + 4 * k * n for k in range(13)]
    rest = base + 4 * 13 * n
    if nb:
        ptrs += [rest, rest + 4 * nb * n, rest + 8 * nb * n]
    else:
        ptrs += [None, None, None]
    ptrs.append(rest + 12 * nb * n)
    return Gauss3D(*ptrs, n, nb)


def splat3d_view(slab: DeviceArray, n: int) -> Splat3D:
    return Splat3D(*[slab.ptr + 4 * k * n for k in range(13)], n)


def splat2d_view(slab: DeviceArray, n: int) -> Splat2D:
    return Splat2D(*[slab.ptr + 4 * k * n for k in range(10)], n)
