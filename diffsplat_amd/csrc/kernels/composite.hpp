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
// The backward writes one gradient record per (tile, splat) pair that a pixel
// of the tile reaches, instead of accumulating with atomics: 3 float4 at
// pair_rec[3 p] = {g0..g3}, {g4..g7}, {g8, epoch bits, 0, 0}, the nine sums
// of the splat's Splat2D gradient fields (pos_x, pos_y, cov_xx, cov_xy,
// cov_yy, r, g, b, opacity), tagged with the forward's epoch (*epoch). p is
// the pair's slot before the tile sort: the splat's pairs are the contiguous
// slots of its tile rectangle in row-major order, so p = pair_base[g].x +
// tile_row * pair_base[g].y + tile_col (the rectangle's width in .y, its
// first slot minus the offset of its first tile in .x, modulo 2^32).
// launch_pair_reduce then adds each splat's records into the gradient.
struct CompositeArgs {
    const float4 *rec;
    const uint32_t *flat;
    const int2 *ranges;
    int W, H, ntx, ntiles;
    const MultiInst *mi;
    const uint16_t *tinst;
    const uint32_t *order;
    float4 *pair_rec;
    const uint2 *pair_base;
    const uint32_t *epoch;
    const int2 *order_range;  // with order: the range of tile order[b], at b
};

// pix_per_thread: 1, 2 or 4 pixels per thread (256, 128 or 64 threads per tile).
ds_status launch_composite_fwd(const CompositeArgs &a, int pix_per_thread, float *out_rgb,
                               float *out_T, int32_t *out_last, const float *target,
                               float *loss, float *grad, hipStream_t s);
// Forward with the fused MSE (target required) and the backward of each tile
// in one launch; the backward writes the gradient records like
// launch_composite_bwd.
ds_status launch_composite_fused(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                                 float *out_rgb, float *out_T, int32_t *out_last,
                                 const float *target, float *loss, float *grad_rgb,
                                 hipStream_t s);
ds_status launch_composite_bwd(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                               const float *T, const int32_t *last, const float *grad_rgb,
                               hipStream_t s);
// Adds the gradient records of the current epoch into grad, per splat: slot p
// (p < *nisect, the pairs kept) belongs to splat slot_splat[p], and a
// splat's slots are contiguous. slot_cap bounds *nisect (grid size).
// grad_is_zero: grad was cleared earlier on the stream and nothing else adds
// to it, so each splat's sum is stored.
// Heavy-first dispatch order of the compositing workgroups (tiles bucketed by
// floor(log2 pairs) + 1, empty tiles 0, buckets in descending order; the
// order within a bucket is arbitrary), computed by one workgroup of the pair
// reduction from the tile ranges of the forward it follows, for the next
// forward: inv[t] is the dispatch position of tile t (a permutation; the
// next forward writes the order and each position's range from it).
// ranges == nullptr: nothing to compute. At most kNextOrderMaxTiles tiles.
constexpr int kNextOrderMaxTiles = 4096;
struct NextOrder {
    const int2 *ranges;
    int ntiles;
    uint32_t *inv;
};

__device__ __forceinline__ uint32_t tile_bucket(int2 r) {
    const uint32_t c = (uint32_t)max(r.y - r.x, 0);
    return c ? (uint32_t)min(31, 32 - __clz(c)) : 0u;
}

// Lanes of this wave (among `valid` ones) whose bucket equals this lane's.
__device__ __forceinline__ uint64_t same_bucket(uint32_t b, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 5; ++bit) {
        const bool on = (b >> bit) & 1u;
        const uint64_t bb = __ballot(on);
        peers &= on ? bb : ~bb;
    }
    return peers;
}

// next: also compute the next forward's dispatch order (see NextOrder).
ds_status launch_pair_reduce(const float4 *pair_rec, const uint32_t *slot_splat,
                             const uint32_t *nisect, const uint32_t *epoch, int64_t slot_cap,
                             const ds_splat2d &grad, bool grad_is_zero, hipStream_t s,
                             NextOrder next = NextOrder{nullptr, 0, nullptr});

}  // namespace ds
