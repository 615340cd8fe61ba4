// Front-to-back compositing of sorted splat lists per 16x16 tile, forward and
// backward (launchers; kernels in composite.hip).
#pragma once

#include "../ds_device.hpp"
#include "../ds_internal.hpp"

namespace ds {

constexpr int kTile = 16;
constexpr int kTilePix = kTile * kTile;
constexpr float kMaxAlpha = 0.99f;
constexpr float kAlphaThres = 1.0f / 255.0f;
constexpr float kTThres = 0.0001f;

// Packed per-splat render record written by tile assignment (3 x float4):
//   [0] = {x, y, k a, k b}   [1] = {k c, opacity, 1/det, splat id bits}
//   [2] = {r, g, b, 0}
// conic = inverse 2D covariance: a = cyy/det, b = -cxy/det, c = cxx/det,
// stored scaled by k = -1/(2 ln 2), so that the falloff exp(-d^2/2) is
// exp2 of the scaled quadratic form directly.
constexpr float kFalloffScale = -0.72134752044448170f;  // -0.5 / ln 2

// One fit instance of a multi-instance rasterizer: its image, its run of the
// global tile ids and of the concatenated pixel arrays, 1 / (W H) for its loss.
struct MultiInst {
    int W, H, ntx, nty;
    int tile_base, pix_base;
    float inv_npix;
    int pad;
};

// mi / tinst are NULL for one instance (W, H, ntx describe it); otherwise
// tile t belongs to instance tinst[t] with the layout mi[tinst[t]].
// Workgroup b composites tile order[b] (NULL: tile b).
struct CompositeArgs {
    const float4 *rec;
    const uint32_t *flat;
    const int2 *ranges;
    int W, H, ntx, ntiles;
    const MultiInst *mi;
    const uint16_t *tinst;
    const uint32_t *order;
};

// pix_per_thread: 1, 2 or 4 pixels per thread (256, 128 or 64 threads per tile).
ds_status launch_composite_fwd(const CompositeArgs &a, int pix_per_thread, float *out_rgb,
                               float *out_T, int32_t *out_last, const float *target,
                               float *loss, float *grad, hipStream_t s);
// Forward with the fused MSE (target required) and the backward of each tile
// in one launch; the backward accumulates into grad like launch_composite_bwd.
ds_status launch_composite_fused(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                                 float *out_rgb, float *out_T, int32_t *out_last,
                                 const float *target, float *loss, float *grad_rgb,
                                 const ds_splat2d &grad, hipStream_t s);
ds_status launch_composite_bwd(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                               const float *T, const int32_t *last, const float *grad_rgb,
                               const ds_splat2d &grad, hipStream_t s);

}  // namespace ds
