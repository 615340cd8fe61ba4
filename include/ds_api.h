/*
 * ds_api.h -- C-ABI of the MI355X (gfx950) Gaussian-splatting engine.
 *
 * Every entry point takes plain pointers and sizes; device pointers are
 * caller-owned unless an object (ds_raster, ds_fitter, ds_renderer) says
 * otherwise.  Every compute call takes a HIP stream (NULL = the null stream)
 * and is asynchronous with respect to the host unless its comment says it
 * synchronises.  Errors are returned as ds_status; ds_last_error() gives the
 * detail of the last failing call on the calling thread.
 *
 * The host-side C++ layer classes (diffsplat_amd/csrc/host/) keep the
 * reference's class names and method semantics on top of these functions:
 *   GaussActivLayer      src/layers/gauss_activ_layer.{h,cu}
 *   PerspProjectLayer    src/layers/persp_project_layer.{h,cu}
 *   TileRasterizeLayer   src/layers/tile_rasterize_layer.{h,cu}
 *   MSELossLayer         src/layers/mse_loss_layer.{h,cu}
 *   Adam                 src/optimizers/adam.{h,cu}
 *   ImageFitter          src/pipelines/image_fitter.{h,cpp}
 *   SplatRenderer        src/pipelines/splat_renderer.{h,cpp}
 */
#ifndef DS_API_H
#define DS_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t, passed as an opaque pointer so that this header needs no HIP headers. */
typedef void *ds_stream_t;

/* Gradient convention of every backward entry point.
 * DS_GRAD_EXACT: the true derivative of the forward functions (agrees with
 *   finite differences at every field, SH degree and camera).
 * DS_GRAD_REFERENCE: the reference's expressions, which differ from the true
 *   derivative in three places:
 *   - compositing dL/dalpha = T_before * (c - colour behind) with the colour
 *     behind not divided by T after the splat
 *     (src/layers/tile_rasterize_layer.cu:491-495);
 *   - the position gradient of the projection omits the dependence of the
 *     Jacobian J in J Sigma J^T on position
 *     (src/layers/persp_project_layer.cu:254-260);
 *   - the SH view direction is treated as constant, so position receives no
 *     gradient through the colour (src/layers/gauss_activ_layer.cu:182-184). */
#define DS_GRAD_EXACT 0
#define DS_GRAD_REFERENCE 1

typedef enum ds_status {
    DS_OK = 0,
    DS_ERR_INVALID_ARG = 1,   /* bad pointer, size or configuration            */
    DS_ERR_HIP = 2,           /* a HIP runtime call failed                      */
    DS_ERR_OUT_OF_MEMORY = 3, /* device allocation failed                       */
    DS_ERR_CAPACITY = 4,      /* a fixed-capacity buffer would overflow         */
    DS_ERR_STATE = 5          /* call order violated (e.g. bwd before fwd)      */
} ds_status;

/* ------------------------------------------------------------------ */
/* Library / runtime helpers                                           */
/* ------------------------------------------------------------------ */

const char *ds_status_string(ds_status s);
/* Detail message of the last failing call on this thread ("" if none). */
const char *ds_last_error(void);
/* Library version string, and the gfx target it was compiled for. */
const char *ds_version(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int ds_device_count(void);
/* Name of device 0 ("" without a GPU). */
ds_status ds_device_name(char *buf, size_t len);

/* Device memory, for callers without their own allocator (tests, benches). */
ds_status ds_malloc(void **dptr, size_t bytes);
ds_status ds_free(void *dptr);
ds_status ds_memcpy_htod(void *dst, const void *src, size_t bytes);
ds_status ds_memcpy_dtoh(void *dst, const void *src, size_t bytes);
ds_status ds_memcpy_dtod(void *dst, const void *src, size_t bytes, ds_stream_t s);
/* Fills `count` 32-bit words with `value` using a kernel (safe under graph capture). */
ds_status ds_fill_u32(void *dptr, uint32_t value, size_t count, ds_stream_t s);
ds_status ds_synchronize(void);
ds_status ds_stream_create(ds_stream_t *s);
ds_status ds_stream_destroy(ds_stream_t s);
ds_status ds_stream_synchronize(ds_stream_t s);

/* HIP events, for timing launches on a given stream. */
typedef void *ds_event_t;
ds_status ds_event_create(ds_event_t *e);
ds_status ds_event_destroy(ds_event_t e);
ds_status ds_event_record(ds_event_t e, ds_stream_t s);
ds_status ds_event_synchronize(ds_event_t e);
ds_status ds_event_elapsed_ms(ds_event_t start, ds_event_t end, float *ms);

/* ------------------------------------------------------------------ */
/* SoA views (one device float array per field, count elements each)  */
/* ------------------------------------------------------------------ */

/* Raw trainable parameters and their gradients.
 * Replaces Gaussian3DParams / Gaussian3DGrads (src/types/gaussian3d.h:34-120).
 * sh_rest_{r,g,b} are band-major: element (band, i) is at [band * count + i];
 * they may be NULL when sh_num_bands == 0. */
typedef struct ds_gauss3d {
    float *pos_x, *pos_y, *pos_z;
    float *scale_x, *scale_y, *scale_z;        /* log-scale            */
    float *rot_w, *rot_x, *rot_y, *rot_z;      /* raw quaternion       */
    float *sh_dc_r, *sh_dc_g, *sh_dc_b;
    float *sh_rest_r, *sh_rest_g, *sh_rest_b;
    float *logit_opacity;
    int32_t count;
    int32_t sh_num_bands;                      /* 0, 3, 8 or 15        */
} ds_gauss3d;

/* World-space activated splats. Replaces Splat3DParams / Splat3DGrads
 * (src/types/splat3d.h:14-51). Covariance is the upper triangle. */
typedef struct ds_splat3d {
    float *pos_x, *pos_y, *pos_z;
    float *cov_xx, *cov_xy, *cov_xz, *cov_yy, *cov_yz, *cov_zz;
    float *color_r, *color_g, *color_b;
    float *opacity;
    int32_t count;
} ds_splat3d;

/* NDC-space splats. Replaces Splat2DParams / Splat2DGrads
 * (src/types/splat2d.h:13-46). */
typedef struct ds_splat2d {
    float *pos_x, *pos_y, *pos_z;
    float *cov_xx, *cov_xy, *cov_yy;
    float *color_r, *color_g, *color_b;
    float *opacity;
    int32_t count;
} ds_splat2d;

/* ------------------------------------------------------------------ */
/* Activation: log-scale/quaternion -> covariance, SH -> colour,       */
/* logit -> opacity.                                                   */
/* ------------------------------------------------------------------ */

/* GaussActivLayer::forward, covForwardKernel (src/layers/gauss_activ_layer.cu:37-170, 411-440).
 * sh_degree in [0,3]; in->sh_num_bands must cover it. cam_pos is the
 * world-space camera position used for the SH view direction. */
ds_status ds_activ_fwd(const ds_gauss3d *in, const ds_splat3d *out, int sh_degree,
                       const float cam_pos[3], ds_stream_t s);

/* GaussActivLayer::backward, covBackwardKernel (src/layers/gauss_activ_layer.cu:198-407, 442-476).
 * `out` is the forward output (clamp gate and sigmoid use it); grad_out holds
 * dL/dSplat3D with off-diagonal covariance gradients in the full-matrix
 * convention; grad_in receives dL/dparams (overwritten, not accumulated). */
ds_status ds_activ_bwd(const ds_gauss3d *in, const ds_splat3d *out, const ds_splat3d *grad_out,
                       const ds_gauss3d *grad_in, int sh_degree, const float cam_pos[3],
                       int grad_mode, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* Perspective projection world -> NDC                                 */
/* ------------------------------------------------------------------ */

/* PerspProjectLayer::setCamera + forward (src/layers/persp_project_layer.cu:40-149, 269-274, 279-292).
 * pv is P*V, column-major (element (row, col) at pv[col*4+row]), passed by
 * value on every call. Splats with clip_w <= 0 get pos_z = FLT_MAX and their
 * other outputs are left unwritten, as in the reference. */
ds_status ds_project_fwd(const ds_splat3d *in, const ds_splat2d *out, const float pv[16],
                         ds_stream_t s);

/* PerspProjectLayer::backward (src/layers/persp_project_layer.cu:173-267, 294-317). */
ds_status ds_project_bwd(const ds_splat3d *in, const ds_splat2d *grad_out,
                         const ds_splat3d *grad_in, const float pv[16], int grad_mode,
                         ds_stream_t s);

/* Fused activation + projection (no Splat3D round trip through HBM).
 * Same results as ds_activ_fwd followed by ds_project_fwd. */
ds_status ds_activ_project_fwd(const ds_gauss3d *in, const ds_splat2d *out, int sh_degree,
                               const float cam_pos[3], const float pv[16], ds_stream_t s);

/* Fused backward: dL/dSplat2D -> dL/dparams. Same results as ds_project_bwd
 * followed by ds_activ_bwd (the clamp gate and sigmoid derivative are
 * recomputed from the parameters instead of read from a saved Splat3D). */
ds_status ds_activ_project_bwd(const ds_gauss3d *in, const ds_splat2d *grad_out,
                               const ds_gauss3d *grad_in, int sh_degree,
                               const float cam_pos[3], const float pv[16], int grad_mode,
                               ds_stream_t s);

/* Fused forward / backward for many fit instances in one launch (no
 * counterpart in the reference): splat i is seen by the camera of instance
 * splat_inst[i] (device uint16 array, e.g. ds_raster_splat_instances), whose
 * record in cam_tab (device, 20 floats per instance) is the 16 PV floats
 * (column-major, as pv above) then cam_pos x, y, z and one unused float.
 * Per splat the results equal ds_activ_project_fwd / _bwd with that camera. */
ds_status ds_activ_project_fwd_multi(const ds_gauss3d *in, const ds_splat2d *out,
                                     int sh_degree, const float *cam_tab,
                                     const uint16_t *splat_inst, ds_stream_t s);
ds_status ds_activ_project_bwd_multi(const ds_gauss3d *in, const ds_splat2d *grad_out,
                                     const ds_gauss3d *grad_in, int sh_degree,
                                     const float *cam_tab, const uint16_t *splat_inst,
                                     int grad_mode, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* Tile rasterizer                                                     */
/* ------------------------------------------------------------------ */

typedef struct ds_raster ds_raster;

typedef struct ds_raster_config {
    int32_t width, height;     /* image size in pixels                                     */
    int32_t count_cap;         /* maximum splat count per call (< 2^30)                    */
    int64_t isect_cap;         /* maximum (tile, splat) pairs (< 2^30); 0 = 1<<25 as the
                                  reference                                                */
    int32_t keep_keys;         /* 1: also materialise the sorted 64-bit (tile|depth) keys  */
    int32_t grad_mode;         /* DS_GRAD_EXACT (0) or DS_GRAD_REFERENCE (1)               */
} ds_raster_config;

/* Device pointers into the rasterizer's own buffers, valid until the next
 * ds_raster_fwd / resize / destroy on the same object. */
typedef struct ds_raster_view {
    const uint64_t *isect_keys_sorted; /* [n_isects] or NULL unless keep_keys               */
    const uint32_t *flatten_ids;       /* [n_isects] splat index per sorted pair             */
    const int32_t *tile_ranges;        /* [num_tiles][2] {start, end}; {0,0} for empty      */
    const float *transmittance;        /* [H*W] final T per pixel                            */
    const int32_t *last_ids;           /* [H*W] sorted index of last contributor             */
    const uint32_t *n_isects;          /* device scalar: pairs kept (<= isect_cap)           */
    const uint32_t *n_isects_wanted;   /* device scalar: pairs requested before the cap      */
    const uint32_t *visible_count;     /* device scalar                                      */
    int32_t num_tiles_x, num_tiles_y;
    int32_t width, height;
    const uint32_t *tile_order;        /* [num_tiles] compositing dispatch order of the last
                                          forward, or NULL when it was row-major          */
    const uint32_t *lookback_error;    /* device word, nonzero when a sort look-back gave up
                                          in some forward since the last report; sticky
                                          across forwards and graph replays, cleared by
                                          the call that reports it (ds_raster_counts, the
                                          fitters' loss / synchronize, ...)               */
} ds_raster_view;

/* TileRasterizeLayer::allocate (src/layers/tile_rasterize_layer.cu:591-616). */
ds_status ds_raster_create(const ds_raster_config *cfg, ds_raster **out);
ds_status ds_raster_destroy(ds_raster *r);
/* TileRasterizeLayer::resize (src/layers/tile_rasterize_layer.cu:618-633). */
ds_status ds_raster_resize(ds_raster *r, int width, int height);

/* Several independent images in one rasterizer (no counterpart in the
 * reference, which rasterizes one image per TileRasterizeLayer): instance k
 * owns the k-th run of `count` splats of the Splat2D input, its own
 * width x height image and 16x16 tiles. One ds_raster_fwd / ds_raster_bwd
 * then runs every phase once for all instances: tile ids are global (the
 * instances' tiles one after another, each row-major), the sorted pairs and
 * tile ranges cover all instances, and the pixel arrays (image, target,
 * dL/dimage, transmittance, last ids) are the instances' HWC images one after
 * another. Within an instance every result equals that of a one-instance
 * rasterizer on its splats (sorted splat ids offset by the instance's first
 * splat). With a target, loss_dev receives one loss per instance. The pair
 * cap is shared: pairs beyond it are dropped farthest depth first over all
 * instances. The one-pass tile sort is not used. Not resizable. */
typedef struct ds_raster_instance {
    int32_t width, height;
    int32_t count;
} ds_raster_instance;
ds_status ds_raster_create_multi(const ds_raster_instance *inst, int n_inst, int64_t isect_cap,
                                 int keep_keys, int grad_mode, ds_raster **out);
/* Device array [total splats] of each splat's instance (uint16), valid for the
 * rasterizer's life. */
ds_status ds_raster_splat_instances(ds_raster *r, const uint16_t **dev);

/* TileRasterizeLayer::forward (src/layers/tile_rasterize_layer.cu:637-708):
 * tile assign -> depth/tile radix sort -> tile ranges -> front-to-back
 * compositing. out_rgb is [H*W*3] HWC. No host synchronisation.
 * If target_rgb is non-NULL the MSE loss of the reference's MSELossLayer
 * (sum over channels / (H*W)) is written to *loss_dev and its gradient to
 * grad_rgb (both device pointers) in the same pass; pass NULLs otherwise. */
ds_status ds_raster_fwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                        const float *target_rgb, float *loss_dev, float *grad_rgb,
                        ds_stream_t s);

/* TileRasterizeLayer::forward followed by MSELossLayer and TileRasterizeLayer::backward
 * (src/layers/tile_rasterize_layer.cu:637-728, src/layers/mse_loss_layer.cu:22-67)
 * in one call: the forward with the fused MSE as ds_raster_fwd, then dL/dSplat2D
 * accumulated into grad_in as ds_raster_bwd (zero it first). When the forward and
 * backward launch shapes are equal, each tile's forward and backward run in one
 * compositing kernel, which keeps the per-pixel state in registers and skips in the
 * backward the (wave, splat) pairs the forward found unreached. The image,
 * transmittance, last ids, loss and dL/dimage are written as by ds_raster_fwd, and
 * ds_raster_bwd may still be called afterwards. No counterpart in the reference. */
ds_status ds_raster_fwd_bwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                            const float *target_rgb, float *loss_dev, float *grad_rgb,
                            const ds_splat2d *grad_in, ds_stream_t s);

/* The compositing of ds_raster_fwd_bwd alone (one fused kernel), on the sorted
 * pairs and tile ranges of the last forward: the splats must be those of that
 * forward. Needs equal forward and backward launch shapes. Zeroes the loss
 * (one float per instance), writes the image, transmittance, last ids and
 * dL/dimage, accumulates into grad_in. For timing and for re-differentiating
 * against another target without sorting again. No counterpart in the
 * reference. */
ds_status ds_raster_composite_fwd_bwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                                      const float *target_rgb, float *loss_dev, float *grad_rgb,
                                      const ds_splat2d *grad_in, ds_stream_t s);

/* Frees the backward's buffers (a 48-byte gradient record per pair slot of
 * isect_cap, 8 bytes per splat) of a rasterizer used for forward passes only;
 * every backward entry point then fails. No counterpart in the reference. */
ds_status ds_raster_release_backward(ds_raster *r);

/* The compositing forward of ds_raster_fwd alone (no MSE), on the sorted pairs
 * and tile ranges of the last forward: the splats must be those of that
 * forward. Writes the image, transmittance and last ids. For timing the
 * compositing apart from the sort. No counterpart in the reference. */
ds_status ds_raster_composite_fwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                                  ds_stream_t s);

/* TileRasterizeLayer::backward (src/layers/tile_rasterize_layer.cu:710-728).
 * Gradients are ACCUMULATED into grad_in, as in the reference: the caller
 * zeroes them first (ds_splat2d_zero). grad_in->pos_z is not written. The
 * reference adds per (tile, splat) with float atomics; here each (tile, splat)
 * pair that the tile's pixels reach gets one gradient record, and a per-splat
 * pass adds a splat's records in slot order (no atomics on grad_in). */
ds_status ds_raster_bwd(ds_raster *r, const ds_splat2d *in, const float *grad_rgb,
                        const ds_splat2d *grad_in, ds_stream_t s);

/* Zeroes all ten arrays of a Splat2D view with fill kernels. */
ds_status ds_splat2d_zero(const ds_splat2d *v, ds_stream_t s);

/* Reads back n_isects and visible_count (synchronises the stream). Replaces
 * TileRasterizeLayer::getVisibleCount (src/layers/tile_rasterize_layer.cu:582-587).
 * n_isects_wanted saturates at 2^30 - 1. Returns DS_ERR_STATE if the last
 * forward's sort could not complete (its output is then not valid). */
ds_status ds_raster_counts(ds_raster *r, ds_stream_t s, uint32_t *n_isects,
                           uint32_t *visible_count, uint32_t *n_isects_wanted);

ds_status ds_raster_get_view(ds_raster *r, ds_raster_view *out);

/* Launch-shape knobs of the compositing kernels (results do not depend on
 * them): pixels of a 16x16 tile per thread in {1, 2, 4} (256, 128 or 64
 * threads per tile), for the forward and the backward separately.
 * Defaults: forward 1, backward 1. */
ds_status ds_raster_set_tuning(ds_raster *r, int fwd_pix_per_thread, int bwd_pix_per_thread);
ds_status ds_raster_get_tuning(ds_raster *r, int *fwd_pix_per_thread, int *bwd_pix_per_thread);

/* Dispatch order of the compositing workgroups (results do not depend on
 * it; no counterpart in the reference, which launches tiles row-major):
 * DS_TILE_ORDER_HEAVY_FIRST (default) starts the tiles with the most pairs
 * first (power-of-two buckets of the pair count, computed on the device
 * after the sort), so the longest tiles do not end the grid late;
 * DS_TILE_ORDER_ROWS is row-major. The backward uses the order of its
 * forward. */
#define DS_TILE_ORDER_ROWS 0
#define DS_TILE_ORDER_HEAVY_FIRST 1
ds_status ds_raster_set_tile_order(ds_raster *r, int order);
ds_status ds_raster_get_tile_order(ds_raster *r, int *order);

/* How the forward sorts the splats by depth before emitting their pairs
 * (results do not depend on it; the reference sorts the 64-bit pair keys in
 * one CUB call, tile_rasterize_layer.cu:679-688):
 * DS_DEPTH_SORT_ONESWEEP, four 8-bit passes of one kernel each with chained
 * look-back (count_cap up to 131072; the default there);
 * DS_DEPTH_SORT_RADIX8, four 8-bit passes of histogram, scan and scatter
 * kernels (the default above 131072);
 * DS_DEPTH_SORT_RADIX11, three 11-bit passes of the same kernels (2048
 * digits per chunk of 2048 keys; count_cap up to 2^24).
 * The tile pass after it keeps 8-bit digits. */
#define DS_DEPTH_SORT_ONESWEEP 0
#define DS_DEPTH_SORT_RADIX8 1
#define DS_DEPTH_SORT_RADIX11 2
ds_status ds_raster_set_depth_sort(ds_raster *r, int mode);
ds_status ds_raster_get_depth_sort(ds_raster *r, int *mode);
/* Times ds_raster_sort on `in` (reps runs after one warm-up, HIP events on s,
 * synchronises) with every mode the raster takes and keeps the fastest (the
 * current one unless another is at least 3% faster). mode (optional): the
 * kept mode; ms (optional, 3 floats): ms per sort for each mode, -1 for a
 * mode the raster does not take. The fitters' and the renderer's autotune
 * calls this. */
ds_status ds_raster_autotune_depth_sort(ds_raster *r, const ds_splat2d *in, int reps,
                                        ds_stream_t s, int *mode, float *ms);
/* The times of the last ds_raster_autotune_depth_sort (3 floats, -1: none). */
ds_status ds_raster_depth_sort_times(ds_raster *r, float *ms);

/* The forward up to compositing: tile assignment, depth and tile sorts,
 * tile ranges and dispatch order, for the splats `in` (no image written).
 * For timing the sort apart from the compositing; a later ds_raster_bwd or
 * ds_raster_composite_* call sees these sorted pairs. No counterpart in the
 * reference. */
ds_status ds_raster_sort(ds_raster *r, const ds_splat2d *in, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* Stand-alone sort (the radix sort inside ds_raster_fwd)              */
/* ------------------------------------------------------------------ */

/* Stable ascending sort of n (u64 key, u32 value) pairs on the low
 * `key_bits` bits, as cub::DeviceRadixSort::SortPairs does at
 * src/layers/tile_rasterize_layer.cu:667-683. Allocates its own scratch
 * (synchronous). */
ds_status ds_sort_pairs_u64(const uint64_t *keys_in, const uint32_t *vals_in,
                            uint64_t *keys_out, uint32_t *vals_out, int64_t n, int key_bits,
                            ds_stream_t s);

/* The same sort without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its scratch
 * is a fixed device array of about 12.6 MiB per GPU, held until the process
 * ends; n is at most 1048576 (2^20) pairs, larger n returns DS_ERR_CAPACITY.
 * Asynchronous on s (unlike ds_sort_pairs_u64). Every call on one GPU shares
 * that GPU's scratch: calls that overlap in time on one GPU race. Output
 * equals ds_sort_pairs_u64's bit for bit. */
ds_status ds_sort_pairs_u64_noalloc(const uint64_t *keys_in, const uint32_t *vals_in,
                                    uint64_t *keys_out, uint32_t *vals_out, int64_t n,
                                    int key_bits, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* MSE loss                                                            */
/* ------------------------------------------------------------------ */

/* MSELossLayer::forward (src/layers/mse_loss_layer.cu:22-42, 84-91):
 * *loss_dev = sum_px sum_c (p - t)^2 / num_pixels. */
ds_status ds_mse_fwd(const float *pixels, const float *target, int64_t num_pixels,
                     float *loss_dev, ds_stream_t s);
/* MSELossLayer::backward (src/layers/mse_loss_layer.cu:54-67, 93-98): grad = 2 (p - t) / num_pixels. */
ds_status ds_mse_bwd(const float *pixels, const float *target, int64_t num_pixels,
                     float *grad, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* Adam                                                                */
/* ------------------------------------------------------------------ */

#define DS_ADAM_MAX_GROUPS 32

typedef struct ds_adam_group {
    float *param;
    const float *grad;
    float *m;
    float *v;
    int64_t n;
    float lr;
    /* Exponential schedule (not in the reference): step t uses
     * lr * decay^(t-1). 0 or 1 keeps lr constant, as the reference does. */
    float decay;
} ds_adam_group;

typedef struct ds_adam_hparams {
    float beta1, beta2, epsilon;   /* reference defaults 0.9, 0.999, 1e-8 (src/optimizers/adam.h:5-9) */
} ds_adam_hparams;

/* Adam::step (src/optimizers/adam.cu:12-37, 59-73) for all groups in ONE
 * launch. `step` is the 1-based step number t used for the bias
 * corrections 1/(1-beta^t). */
ds_status ds_adam_step(const ds_adam_group *groups, int n_groups, const ds_adam_hparams *hp,
                       int32_t step, ds_stream_t s);

/* Same, with the step counter in device memory: *step_dev is incremented
 * on the device and the new value used, so a captured graph can replay it. */
ds_status ds_adam_step_dev(const ds_adam_group *groups, int n_groups,
                           const ds_adam_hparams *hp, int32_t *step_dev, ds_stream_t s);

/* ------------------------------------------------------------------ */
/* Image fitter (ImageFitter, src/pipelines/image_fitter.{h,cpp})      */
/* ------------------------------------------------------------------ */

typedef struct ds_fitter ds_fitter;

typedef struct ds_fitter_config {
    int32_t width, height;
    int32_t count;
    int32_t sh_degree;
    float view[16], proj[16];      /* column-major, OpenGL conventions                */
    float cam_pos[3];
    float lr_pos, lr_scale, lr_rot, lr_sh_dc, lr_sh_rest, lr_opacity;
    float beta1, beta2, epsilon;
    int64_t isect_cap;             /* 0 = 1<<25                                        */
    int32_t fused;                 /* 1: fused activation+projection and fused MSE     */
    int32_t grad_mode;             /* DS_GRAD_EXACT or DS_GRAD_REFERENCE               */
    float lr_pos_decay;            /* position lr schedule per step (ds_adam_group.decay);
                                      1 = constant as the reference                    */
} ds_fitter_config;

/* Fills cfg with the reference's defaults (src/pipelines/image_fitter.cpp:18-34,
 * src/optimizers/gaussian3d_adam.h:11-19): ortho pixel-unit camera, 3DGS LRs. */
void ds_fitter_default_config(int width, int height, int count, ds_fitter_config *cfg);

ds_status ds_fitter_create(const ds_fitter_config *cfg, ds_fitter **out);
ds_status ds_fitter_destroy(ds_fitter *f);
/* Uploads the target image, host [H*W*3] HWC floats. */
ds_status ds_fitter_set_target(ds_fitter *f, const float *host_rgb);
/* Parameter slab layout (host, field-major, n = count, b = bands):
 * pos_x,pos_y,pos_z, scale_x..z, rot_w..z, sh_dc_r..b  (13 x n),
 * sh_rest_r, sh_rest_g, sh_rest_b (3 x b*n), logit_opacity (n).
 * Uploading resets Adam state and the step counter. */
ds_status ds_fitter_upload_params(ds_fitter *f, const float *host_slab);
ds_status ds_fitter_download_params(ds_fitter *f, float *host_slab);
/* Same layout, gradients of the last step. */
ds_status ds_fitter_download_grads(ds_fitter *f, float *host_slab);
/* Adam first and second moments (Adam::step state, src/optimizers/adam.cu:12-73)
 * in the same slab layout; either pointer may be NULL (synchronises). */
ds_status ds_fitter_download_moments(ds_fitter *f, float *host_m, float *host_v);
/* ImageFitter::step (src/pipelines/image_fitter.cpp:103-112), n_steps times,
 * asynchronously on the fitter's stream. */
ds_status ds_fitter_step(ds_fitter *f, int n_steps);
/* Forward only (render + loss) with the current parameters. */
ds_status ds_fitter_forward(ds_fitter *f);
/* Loss of the last forward (synchronises). This call, ds_fitter_synchronize
 * and ds_fitter_download_image return DS_ERR_STATE (and clear the flag) when
 * a sort look-back of any step since the last report gave up waiting
 * (ds_raster_view.lookback_error); *loss is still written. */
ds_status ds_fitter_loss(ds_fitter *f, float *loss);
ds_status ds_fitter_synchronize(ds_fitter *f);
/* Rendered image of the last forward, host [H*W*3] (synchronises). */
ds_status ds_fitter_download_image(ds_fitter *f, float *host_rgb);
ds_status ds_fitter_get_raster(ds_fitter *f, ds_raster **r);
ds_status ds_fitter_get_stream(ds_fitter *f, ds_stream_t *s);
/* 1: capture one step into a HIP graph on first use and replay it. Every
 * call drops a captured step, so call it again after changing the
 * rasterizer's launch shapes or dispatch order directly (ds_raster_set_*). */
ds_status ds_fitter_set_graph(ds_fitter *f, int enable);
/* Graph mode: captures (records, does not run) the graphs that
 * ds_fitter_step(f, n_steps) replays, so that the next such call launches
 * them without capturing or instantiating; a no-op without graphs. For
 * callers that time or latency-bound runs (a setup call, like an
 * allocation). No counterpart in the reference, which launches layer by
 * layer. */
ds_status ds_fitter_prepare(ds_fitter *f, int n_steps);
/* Autotuner: times every compositing launch shape on the current parameters
 * (forward and backward only; parameters and Adam state are unchanged; `reps`
 * timed runs each) and switches to the fastest forward and backward shapes
 * when they beat the current ones by at least 5%. Synchronises. Any of the
 * outputs may be NULL. */
ds_status ds_fitter_autotune(ds_fitter *f, int reps, int *fwd_pix_per_thread,
                             int *bwd_pix_per_thread, float *fwd_ms, float *bwd_ms);
ds_status ds_fitter_step_count(ds_fitter *f, int32_t *step);
/* Device views of the fitter's own buffers (valid for the fitter's life):
 * the parameters, the Splat2D the rasterizer consumes, its gradient, and
 * dL/d(image) [H*W*3]. Any pointer may be NULL. */
ds_status ds_fitter_views(ds_fitter *f, ds_gauss3d *params, ds_splat2d *splat2d,
                          ds_splat2d *grad_splat2d, float **grad_rgb);

/* Forward and backward of one step without the Adam update (zeroGrad,
 * forward, backward), asynchronous: the parameter gradients are then in the
 * gradient slab. For callers that own the optimisation, such as the learned
 * initializer's training. No counterpart in the reference. */
ds_status ds_fitter_forward_backward(ds_fitter *f);
/* Device pointers to the parameter slab and the parameter-gradient slab
 * (layout of ds_fitter_upload_params), valid for the fitter's life. Writes
 * to the parameter slab must be ordered with the fitter's stream. */
ds_status ds_fitter_param_slabs(ds_fitter *f, float **params_dev, const float **grads_dev);
/* Target image from device memory [H*W*3], copied on the fitter's stream. */
ds_status ds_fitter_set_target_device(ds_fitter *f, const float *dev_rgb);
/* Adam moments and step counter to zero; the parameters are kept
 * (synchronises). */
ds_status ds_fitter_reset_optimizer(ds_fitter *f);

/* ------------------------------------------------------------------ */
/* Adaptive density control (clone / split / prune). The reference       */
/* lists it as a TODO (README.md:10); the rules are those of Kerbl et   */
/* al., SIGGRAPH 2023, sec. 5.2.                                         */
/* ------------------------------------------------------------------ */

#define DS_DENSIFY_PRUNE 0
#define DS_DENSIFY_KEEP 1
#define DS_DENSIFY_CLONE 2
#define DS_DENSIFY_SPLIT 3

typedef struct ds_densify_config {
    float grad_threshold;   /* mean |dL/d(NDC centre)| over the steps a splat was visible
                               at or above which it is cloned or split (3DGS: 2e-4)      */
    float percent_dense;    /* clone when the largest scale <= percent_dense * extent,
                               else split (3DGS: 0.01)                                   */
    float scene_extent;     /* world units; <= 0: the largest distance of a splat centre
                               from their mean (as 3DGS's camera extent)                 */
    float min_opacity;      /* prune below this opacity (3DGS: 0.005)                    */
    float max_scale_frac;   /* prune when the largest scale > max_scale_frac * extent;
                               0 = off (3DGS: 0.1, after its first opacity reset)       */
    float split_scale_div;  /* split samples' scale divisor (3DGS: 1.6 = 0.8 * 2)        */
    uint64_t seed;          /* split samples                                             */
} ds_densify_config;

typedef struct ds_densify_result {
    int32_t count_before, count_after;
    int32_t pruned, cloned, split;   /* splats of each action (a split adds 2)        */
} ds_densify_result;

/* 3DGS's defaults (above). */
void ds_densify_default_config(ds_densify_config *cfg);
/* 1: every later fit step also accumulates, per splat visible in that step,
 * |dL/d(NDC centre)| and one visit (one more launch per step); 0: stops.
 * Either way the statistics start from zero. */
ds_status ds_fitter_set_densify_stats(ds_fitter *f, int enable);
/* The accumulated statistics, host [count] each (synchronises). */
ds_status ds_fitter_densify_stats(ds_fitter *f, float *host_grad_sum, uint32_t *host_visits);
/* One densification from the accumulated statistics: classifies every splat
 * (per-splat action codes into host_actions [count before] unless NULL),
 * then compacts the parameter slab and the Adam moments on the device
 * (survivors in order with their moments, then clones, then split samples
 * with zero moments). The splat count changes; the step counter, learning
 * rates, target, camera and launch shapes are kept, the statistics restart
 * from zero and a captured step is captured again. Synchronises. */
ds_status ds_fitter_densify(ds_fitter *f, const ds_densify_config *cfg,
                            ds_densify_result *result, uint8_t *host_actions);
/* Current splat count (changes with ds_fitter_densify). */
ds_status ds_fitter_count(ds_fitter *f, int32_t *count);

/* ------------------------------------------------------------------ */
/* Learned initializer (no counterpart in the reference, which starts    */
/* every fit from random splats, src/utils/splat_utils.h:10-51)           */
/* ------------------------------------------------------------------ */

/* A CNN of six 3x3 convolutions on bf16 MFMA with fp32 accumulation
 * (3 -> 32 /2, 32 -> 32 /2, 32 -> 64 /2, 64 -> 64 /2, 64 -> 64, 64 -> 160;
 * ReLU between) maps a W x H target image to a grid of 16x16-pixel cells
 * with 16 SH-0 splats each ((W/4)*(H/4) splats, one per 4x4-pixel anchor)
 * for the pinhole camera at the origin looking down -z with vertical field
 * of view `fovy` (radians): each splat is a flat ellipse facing the camera,
 * its image position within 16 pixels of its anchor and its depth within
 * 4 +- 1 (which orders the splats), with learned apparent anisotropic
 * scale, in-plane rotation, colour and opacity (10 head values per splat:
 * offset x, y, log-scale major, minor, angle, rgb, opacity, depth). It is
 * trained on (image,
 * fitted splats) pairs by regression of the head values
 * (ds_initnet_train_step_supervised), or end to end through the
 * differentiable renderer (ds_initnet_train_step). */
typedef struct ds_initnet ds_initnet;

/* W and H multiples of 16 in [16, 4096]; weights drawn from `seed`. */
ds_status ds_initnet_create(int width, int height, float fovy, uint64_t seed, ds_initnet **out);
ds_status ds_initnet_destroy(ds_initnet *n);
/* Number of weights and biases (the ds_initnet_get_params layout) and of
 * splats the net produces. Either output may be NULL. */
ds_status ds_initnet_info(ds_initnet *n, int64_t *num_params, int32_t *num_splats);
/* Weights and biases, host floats: per layer W[co][kh][kw][ci] then b[co].
 * Setting them resets the net's optimizer. Both synchronise. */
ds_status ds_initnet_get_params(ds_initnet *n, float *host);
ds_status ds_initnet_set_params(ds_initnet *n, const float *host);
/* Image (device, [H*W*3] in [0,1]) -> SH-0 parameter slab (device, the
 * ds_fitter_upload_params layout for num_splats splats). Asynchronous. */
ds_status ds_initnet_infer(ds_initnet *n, const float *dev_rgb, float *dev_slab, ds_stream_t s);
/* Head values of the last forward, host (synchronises): [cells][160] with
 * cells in row-major order and the 16 splats of a cell in row-major order
 * inside it, 10 values each. */
ds_status ds_initnet_head(ds_initnet *n, float *host_h);
/* The regression targets of a fitted SH-0 slab (device, the
 * ds_fitter_upload_params layout for num_splats splats): the head values
 * (device, the ds_initnet_head layout) whose splats are closest to it, each
 * value through the inverse of its head function (arguments of tanh and
 * sigmoid clamped to 98% of their range): image position and depth; the
 * fitted covariance projected along the view rays as the apparent
 * log-scales (major axis first) and the major axis' angle; colour; opacity.
 * Asynchronous. */
ds_status ds_initnet_head_targets(ds_initnet *n, const float *dev_slab, float *dev_targets,
                                  ds_stream_t s);
/* Weights of the 10 head values in the supervised loss (NULL: the
 * defaults). Non-negative and finite. */
ds_status ds_initnet_set_loss_weights(ds_initnet *n, const float *weights);
/* One supervised training step: forward, mean squared error of the head
 * values against dev_targets (weighted per value; its mean over the splats
 * added to *dev_loss unless NULL), backward,
 * one Adam step (lr). Asynchronous on s. */
ds_status ds_initnet_train_step_supervised(ds_initnet *n, const float *dev_rgb,
                                           const float *dev_targets, float lr, float *dev_loss,
                                           ds_stream_t s);
/* Gradients of all weights and biases (ds_initnet_get_params layout, host)
 * for the image and a given gradient of the output slab (device).
 * Synchronises; leaves the weights unchanged. */
ds_status ds_initnet_grads(ds_initnet *n, const float *dev_rgb, const float *dev_grad_slab,
                           float *host_grads);
/* One training step on one image: inference into the fitter's parameters,
 * the fitter's forward and backward against the same image, backpropagation
 * into the net, one Adam step (lr) on its weights. The fitter (count =
 * num_splats, SH degree 0, same camera and size) is only used to render; its
 * own optimizer is not stepped. Asynchronous on the fitter's stream. */
ds_status ds_initnet_train_step(ds_initnet *n, ds_fitter *f, const float *dev_rgb, float lr);

/* ------------------------------------------------------------------ */
/* Batched fitting: many independent fitters stepped together (no        */
/* counterpart in the reference, which fits one image per process)       */
/* ------------------------------------------------------------------ */

/* Each instance replays its own captured step (graphs are switched on);
 * instances are spread over a few streams, largest measured step first, each
 * to the least-loaded stream, so that small fits overlap on the GPU. The
 * fitters are borrowed: destroy the batch before them. While in a batch a
 * fitter's other calls still work (they run on its batch stream). */
typedef struct ds_batch ds_batch;

ds_status ds_batch_create(ds_fitter *const *fitters, int n, ds_batch **out);
ds_status ds_batch_destroy(ds_batch *b);
/* n_steps steps of every instance. Asynchronous. */
ds_status ds_batch_step(ds_batch *b, int n_steps);
ds_status ds_batch_synchronize(ds_batch *b);
ds_status ds_batch_set_streams(ds_batch *b, int n_streams);
/* Tunes every instance's compositing launch shapes, measures every
 * instance's step time, then times `reps` batch steps with 1, 2, 4 and 8
 * streams and keeps the fastest (it takes real optimisation steps).
 * Synchronises. Either output may be NULL. */
ds_status ds_batch_autotune(ds_batch *b, int reps, int *n_streams, float *ms_per_batch_step);

/* Many fits as one (no counterpart in the reference): every phase of a step
 * is ONE launch over all instances (activation + projection with a per-
 * instance camera table, a multi-instance ds_raster for tile assign, sorts,
 * compositing with the per-instance MSE and its backward, one Adam launch
 * over one parameter slab), and the whole step can be captured into one HIP
 * graph. Per instance a step equals ds_fitter_step with the same camera,
 * learning rates and grad mode, up to the order of float atomics. All
 * instances share the SH degree, the learning rates, the Adam step counter
 * and the pair cap (isect_cap, 0 = 1<<25). */
typedef struct ds_multifit ds_multifit;

typedef struct ds_multifit_instance {
    int32_t width, height;
    int32_t count;
    float view[16], proj[16];      /* column-major, OpenGL conventions */
    float cam_pos[3];
} ds_multifit_instance;

/* cfg supplies sh_degree, the learning rates, lr_pos_decay, beta1/beta2/
 * epsilon and grad_mode (its size, camera, isect_cap and fused fields are
 * not used). */
ds_status ds_multifit_create(const ds_multifit_instance *inst, int n_inst,
                             const ds_fitter_config *cfg, int64_t isect_cap, ds_multifit **out);
ds_status ds_multifit_destroy(ds_multifit *m);
/* Instance k's target [H*W*3] HWC (host). */
ds_status ds_multifit_set_target(ds_multifit *m, int k, const float *host_rgb);
/* Instance k's parameters in the ds_fitter_upload_params layout. Uploading
 * resets the Adam state of every instance and the shared step counter. */
ds_status ds_multifit_upload_params(ds_multifit *m, int k, const float *host_slab);
ds_status ds_multifit_download_params(ds_multifit *m, int k, float *host_slab);
/* Gradients of the last step (same layout). Synchronises. */
ds_status ds_multifit_download_grads(ds_multifit *m, int k, float *host_slab);
/* Rendered image of instance k from the last forward. Synchronises. */
ds_status ds_multifit_download_image(ds_multifit *m, int k, float *host_rgb);
/* n_steps steps of every instance, asynchronous on the multifit's stream. */
ds_status ds_multifit_step(ds_multifit *m, int n_steps);
/* Forward only (render + loss of every instance). */
ds_status ds_multifit_forward(ds_multifit *m);
/* zero grads, forward, backward (no Adam update). */
ds_status ds_multifit_forward_backward(ds_multifit *m);
/* Loss of every instance from the last forward, host [n_inst]. Synchronises. */
/* Both report a sticky look-back give-up as ds_fitter_loss does. */
ds_status ds_multifit_losses(ds_multifit *m, float *losses);
ds_status ds_multifit_synchronize(ds_multifit *m);
/* 1: capture one step into one HIP graph on first use and replay it. Every
 * call drops a captured step (as ds_fitter_set_graph). */
ds_status ds_multifit_set_graph(ds_multifit *m, int enable);
/* As ds_fitter_autotune, over the compositing launches of all instances. */
ds_status ds_multifit_autotune(ds_multifit *m, int reps, int *fwd_pix_per_thread,
                               int *bwd_pix_per_thread, float *fwd_ms, float *bwd_ms);
ds_status ds_multifit_step_count(ds_multifit *m, int32_t *step);
ds_status ds_multifit_get_raster(ds_multifit *m, ds_raster **r);
ds_status ds_multifit_get_stream(ds_multifit *m, ds_stream_t *s);
/* Device views of the multifit's own buffers over all instances, as
 * ds_fitter_views: instance k's splats are elements [sum of the counts of
 * instances < k, + its count) of every field; its pixels follow the pixels
 * of instances < k in dL/d(image). Any pointer may be NULL. */
ds_status ds_multifit_views(ds_multifit *m, ds_gauss3d *params, ds_splat2d *splat2d,
                            ds_splat2d *grad_splat2d, float **grad_rgb);

/* ------------------------------------------------------------------ */
/* Forward-only renderer (SplatRenderer, src/pipelines/splat_renderer.{h,cpp}) */
/* ------------------------------------------------------------------ */

typedef struct ds_renderer ds_renderer;

ds_status ds_renderer_create(int width, int height, int count, int sh_degree, int64_t isect_cap,
                             ds_renderer **out);
ds_status ds_renderer_destroy(ds_renderer *r);
/* Same slab layout as ds_fitter_upload_params. */
ds_status ds_renderer_upload_params(ds_renderer *r, const float *host_slab);
/* SplatRenderer::render (src/pipelines/splat_renderer.cpp:85-91). Async. */
ds_status ds_renderer_render(ds_renderer *r, const float view[16], const float proj[16],
                             const float cam_pos[3]);
/* Autotuner (no counterpart in the reference): re-renders the last camera
 * `reps` times per compositing launch shape, switches to the fastest when it
 * beats the current one by at least 5%. Call after one ds_renderer_render.
 * Synchronises. Either output may be NULL. */
ds_status ds_renderer_autotune(ds_renderer *r, int reps, int *fwd_pix_per_thread,
                               float *ms_per_frame);
ds_status ds_renderer_download_image(ds_renderer *r, float *host_rgb);
ds_status ds_renderer_get_raster(ds_renderer *r, ds_raster **out);
/* Device view of the projected splats of the last frame (the compositing's
 * input), valid for the renderer's life. */
ds_status ds_renderer_views(ds_renderer *r, ds_splat2d *splat2d);
ds_status ds_renderer_get_stream(ds_renderer *r, ds_stream_t *s);
/* This call and ds_renderer_download_image report a sticky look-back give-up
 * as ds_fitter_loss does. */
ds_status ds_renderer_synchronize(ds_renderer *r);

/* ------------------------------------------------------------------ */
/* Scene files (host only)                                              */
/* ------------------------------------------------------------------ */

/* 3DGS PLY: binary little-endian, float vertex properties by name (x y z,
 * scale_0..2, rot_0..3 = w x y z, f_dc_0..2, f_rest_* channel-first, opacity;
 * others are ignored). The file is in the OpenCV convention (y down, z
 * forward), slabs in OpenGL (y up, z backward): y, z and the quaternion's
 * y, z change sign on load and save. SH degree from the f_rest count as in
 * PLYLoader::load (src/io/ply_io.cpp:15-160). ASCII files and non-float
 * vertex properties are rejected (the reference would misread them). */
ds_status ds_ply_info(const char *path, int64_t *count, int32_t *sh_degree);
/* Into a host slab of ds_slab_floats(count, sh_degree) floats (count and
 * sh_degree as ds_ply_info reports). */
ds_status ds_ply_load(const char *path, float *host_slab, int64_t count, int32_t sh_degree);
/* PLYSaver::save (src/io/ply_io.cpp:164-227); property order x y z f_dc
 * f_rest opacity scale rot. */
ds_status ds_ply_save(const char *path, const float *host_slab, int64_t count,
                      int32_t sh_degree);
/* SplatUtils::normalizeScene (src/utils/splat_utils.h:53-93), bit-identical:
 * positions centred on their float centroid and scaled so that the largest
 * |coordinate| is scene_scale; log-scales shifted by log of that factor. */
ds_status ds_normalize_scene(float *host_slab, int64_t count, float scene_scale);
/* SplatRenderer::loadPLY (src/pipelines/splat_renderer.cpp:27-35): load,
 * normalise with scene_scale, create a renderer holding the scene. count and
 * sh_degree (either may be NULL) receive the scene's. */
ds_status ds_renderer_load_ply(const char *path, int width, int height, float scene_scale,
                               int64_t isect_cap, ds_renderer **out, int64_t *count,
                               int32_t *sh_degree);
/* ImageFitter::savePLY (src/pipelines/image_fitter.cpp:114-118). Synchronises. */
ds_status ds_fitter_save_ply(ds_fitter *f, const char *path);

/* Images: PNG (8 or 16 bit, grey / RGB / palette / grey+alpha / RGBA, not
 * interlaced; alpha dropped) or binary PPM (P6, maxval 255). JPEG, which the
 * reference also reads through stb_image, is not supported. */
ds_status ds_image_info(const char *path, int32_t *width, int32_t *height);
/* ImageLoader::load (src/io/image_io.cpp:23-78): the image scaled to fit
 * inside the canvas minus `padding` on every side (scale = min((W-2p)/w,
 * (H-2p)/h), fit size (int)(w*scale) x (int)(h*scale)), centred, black
 * elsewhere, HWC floats value/255 into host_rgb [H*W*3]. The resampling is
 * this library's (area average when shrinking, linear between pixel centres
 * when growing, on 8-bit values), not stb_image_resize2's. src_w, src_h (may
 * be NULL) receive the file's size. */
ds_status ds_image_load(const char *path, int target_w, int target_h, int padding,
                        float *host_rgb, int32_t *src_w, int32_t *src_h);
/* ImageSaver::saveAsPNG (src/io/image_io.cpp:82-93): 8-bit RGB PNG of HWC
 * floats, each byte (uint8)min(max(v * 255, 0), 255). */
ds_status ds_image_save_png(const char *path, const float *host_rgb, int width, int height);
/* Same from device memory (synchronises), as the reference's saveAsPNG takes. */
ds_status ds_image_save_png_device(const char *path, const float *dev_rgb, int width,
                                   int height);

/* ------------------------------------------------------------------ */
/* Synthetic workloads and cameras (host only, deterministic)          */
/* ------------------------------------------------------------------ */

/* Target image: smooth gradient + 3 sinusoids + n_ellipses filled ellipses
 * of uniform colour, [H*W*3] HWC in [0,1]. */
void ds_gen_target_image(uint64_t seed, int width, int height, int n_ellipses, float *out_rgb);

/* Splats inside a pinhole frustum (camera at the origin looking down -z,
 * vertical fov fovy_deg, aspect width/height, depth in [zmin, zmax]):
 * log-scale ~ N(ls_mean, ls_std), quaternion = normalize(N(0,1)^4),
 * logit = logit_opacity, colour ~ U[0,1] (stored as DC SH). Writes the
 * slab layout of ds_fitter_upload_params with sh_degree 0. */
void ds_gen_frustum_splats(uint64_t seed, int count, float fovy_deg, float aspect, float zmin,
                           float zmax, float ls_mean, float ls_std, float logit_opacity,
                           float *slab);

/* Clustered scene: positions from n_clusters anisotropic Gaussian clusters
 * in the unit cube, log-scale ~ N(ls_mean, ls_std), opacity logit ~ N(0,2),
 * SH degree sh_degree with rest coefficients ~ N(0, 0.2) and DC colour
 * N(0.5, 0.2) (stored as DC SH). Slab layout as above. */
void ds_gen_cluster_scene(uint64_t seed, int count, int sh_degree, int n_clusters, float ls_mean,
                          float ls_std, float *slab);

/* The reference fitter's random init semantics (src/utils/splat_utils.h:10-51)
 * on a counter-based generator: pixel-unit box, isotropic log-scale
 * ln2 + U[-0.5,0.5], identity rotation, random DC colour, opacity in [0.6,1]. */
void ds_gen_fitter_init(uint64_t seed, int count, int width, int height, float *slab);

/* OpenGL-convention camera matrices (column-major, clip z in [-1,1]). */
void ds_mat_perspective(float fovy_rad, float aspect, float znear, float zfar, float out[16]);
void ds_mat_ortho(float l, float r, float b, float t, float n, float f, float out[16]);
void ds_mat_look_at(const float eye[3], const float center[3], const float up[3], float out[16]);
void ds_mat_mul(const float a[16], const float b[16], float out[16]);

/* ------------------------------------------------------------------------
 * Interactive cameras: per-frame input -> view/projection for the renderer.
 * Replace ArcballCamera (src/camera/arcball_camera.h:20-71, .cpp:11-144) and
 * FlyCamera (src/camera/fly_camera.h:30-76, .cpp:13-125) behind the Camera
 * interface (src/camera/camera.h:5-16). Host-only; no GPU involved.
 *
 * Both start at eye (0, 0, 3); the arcball looks at the origin, the fly
 * camera down -z. Arcball: left drag orbits (yaw/pitch, pitch within
 * +-(pi/2 - 0.01)), shift + left drag pans eye and target together, scroll
 * dollies along the view direction (distance clamped to [zoom_min_dist,
 * zoom_max_dist]) or, in ortho mode, scales the frustum half-height (>= 0.01).
 * Fly: W/S A/D R/F move along the local forward/right/up axes, Q/E roll, left
 * drag yaws about the local up axis and pitches (clamped as above); shift and
 * ctrl multiply every speed. Mouse deltas are clamped to +-50 pixels, screen y
 * grows downwards, speeds are per second and scale with dt.
 * ---------------------------------------------------------------------- */
#define DS_CAMERA_ARCBALL 0
#define DS_CAMERA_FLY 1

/* Keys the cameras read (the reference's GLFW key queries). */
#define DS_KEY_SHIFT (1u << 0)
#define DS_KEY_CTRL (1u << 1)
#define DS_KEY_W (1u << 2)
#define DS_KEY_A (1u << 3)
#define DS_KEY_S (1u << 4)
#define DS_KEY_D (1u << 5)
#define DS_KEY_R (1u << 6)
#define DS_KEY_F (1u << 7)
#define DS_KEY_Q (1u << 8)
#define DS_KEY_E (1u << 9)

/* One frame of input (src/input/input.h:12-16 and its key set). */
typedef struct ds_camera_input {
    float mouse_dx, mouse_dy; /* cursor movement since the last frame, pixels */
    float scroll;             /* wheel delta since the last frame             */
    int32_t mouse_left;       /* left button held                             */
    uint32_t keys;            /* DS_KEY_* held                                */
} ds_camera_input;

/* Speeds (defaults: arcball_camera.h:40-44, fly_camera.h:44-48). The first
 * five drive the arcball, the rest the fly camera. */
typedef struct ds_camera_speeds {
    float rotate;        /* rad per pixel per second            0.2   */
    float pan;           /* distance fraction per pixel          0.005 */
    float zoom;          /* distance fraction per scroll unit    0.1   */
    float zoom_min_dist; /*                                      0.5   */
    float zoom_max_dist; /*                                      50    */
    float move;          /* world units per second               5     */
    float look;          /* rad per pixel per second             0.2   */
    float roll;          /* rad per second                       0.2   */
    float mult_shift;    /*                                      4     */
    float mult_ctrl;     /*                                      0.2   */
} ds_camera_speeds;

typedef struct ds_camera ds_camera;

void ds_camera_default_speeds(ds_camera_speeds *out);
/* ArcballCamera/FlyCamera(aspect, fovDegrees, nearPlane, farPlane); the
 * reference's defaults are 40 degrees, 0.1, 100. */
ds_status ds_camera_create(int kind, float aspect, float fovy_deg, float znear, float zfar,
                           ds_camera **out);
void ds_camera_destroy(ds_camera *cam);
/* Camera::update (camera.h:10): *changed = 1 if the matrices changed. */
ds_status ds_camera_update(ds_camera *cam, const ds_camera_input *in, float dt, int *changed);
/* Camera::setAspect (camera.h:11). */
ds_status ds_camera_set_aspect(ds_camera *cam, float aspect);
/* Camera::setOrthoMode (camera.h:12): arcball only, a no-op for the fly camera. */
ds_status ds_camera_set_ortho(ds_camera *cam, int enable);
ds_status ds_camera_get_speeds(const ds_camera *cam, ds_camera_speeds *out);
ds_status ds_camera_set_speeds(ds_camera *cam, const ds_camera_speeds *speeds);
/* getViewMatrix / getProjectionMatrix / getPosition (camera.h:14-16); any
 * output may be null. */
ds_status ds_camera_matrices(const ds_camera *cam, float view[16], float proj[16], float eye[3]);

/* Slab size in floats for count splats of sh_degree: (14 + 3*bands) * count. */
int64_t ds_slab_floats(int count, int sh_degree);
int ds_sh_bands(int sh_degree);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* DS_API_H */
