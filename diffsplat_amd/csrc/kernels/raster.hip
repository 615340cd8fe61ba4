// Tile rasterizer: tile assignment, (tile | depth) radix sort, tile ranges,
// front-to-back compositing and its analytic backward.
//
// Reference semantics (src/layers/tile_rasterize_layer.cu):
//   makeKey                 30-38
//   tileAssignKernel        52-144
//   buildTileRangesKernel   148-162
//   compositeFwdKernel      192-309
//   compositeBwdKernel      356-578
//   allocate/resize/forward/backward  591-728
//
// MI355X design:
//  * Tile assignment writes a per-splat tile rectangle, pair count and a
//    packed 48-byte render record (centre, conic, opacity, 1/det, rgb) so
//    compositing gathers one contiguous record per pair instead of nine SoA
//    arrays.
//  * The 64-bit (tile << 32 | depth) sort is factored: every pair of one
//    splat shares the low 32 key bits, so the four LSD passes over the depth
//    bits run over N splats instead of over the pairs; pairs are then emitted
//    in depth order at scanned offsets (deterministic, no atomic slots) and
//    stably sorted on the tile bits only. The result equals a stable sort of
//    the reference's 64-bit keys.
//  * Compositing: one 256-thread workgroup (4 wave64) per 16x16 tile,
//    256-record batches in LDS, tiles dispatched heaviest first (pair count
//    bucketed by powers of two) so that the grid ends evenly. The MSE loss
//    and its gradient can be fused into the forward epilogue.
//  * Backward: wave64 reductions, then LDS float atomics across the four
//    waves, then one 48-byte gradient record per (tile, splat) pair that the
//    tile's pixels reach, stored at the pair's slot before the tile sort
//    (the splat's slots are contiguous) with the forward's epoch as its tag;
//    no global atomics. A per-splat pass then adds each splat's records of
//    this epoch into dL/dSplat2D.
#include <float.h>

#include <algorithm>
#include <vector>

#include "../ds_device.hpp"
#include "../ds_internal.hpp"
#include "composite.hpp"
#include "forward_start.hpp"
#include "sort_scan.hpp"

namespace ds {

namespace {

// Scalars kept on the device: indices into ds_raster::scalars. Those below
// SC_EPOCH and the tickets start each forward at zero (SC_COUNT at n).
// SC_ERR is sticky: no forward clears it, so a give-up in any step of a
// replayed multi-step graph survives until the host reads and clears it
// (raster_take_error).
enum {
    SC_NISECT = 0,
    SC_WANTED = 1,
    SC_VISIBLE = 2,
    SC_COUNT = 3,
    SC_EPOCH = 8,    // look-back epoch: +1 every forward, never 0
    SC_ERR = 9,      // a look-back spin gave up since the host last read it
    SC_TICKET = 16,  // one ticket counter per look-back slot
    SC_NUM = 32
};

constexpr const char *kLookbackError =
    "a sort or scan look-back gave up waiting in a forward since the error was last "
    "reported; the results of that forward (and of the steps after it) are not valid";

// Look-back slots (and digit-histogram rows): the four depth passes.
enum { LB_DEPTH = 0, LB_SLOTS = 8 };
constexpr int kHistRows = 4;  // digit histograms of the depth passes
// Largest splat count sorted with one-kernel passes (64 chunks).
constexpr uint32_t kOneKernelMaxKeys = 64u * kSortChunk;

// Device-scope atomics on one cache line serialise (about 11-19 ns each on
// MI355X), so per-workgroup counts go to replicas that readers sum: the
// digit histograms (kHistReps copies of kHistRows x 256 bins) and the
// visible-splat count (kVisReps words, one per 128-byte line), all in the
// ghist buffer.
constexpr int kHistReps = 8;
constexpr int kHistRepWords = kHistRows * kRadix;
constexpr int kVisReps = 16;
constexpr int kVisStride = 32;
constexpr int kVisOffset = kHistReps * kHistRepWords;
// kHistReps copies of 256 64-bit sums: per top depth byte, the tile count of
// the splats with that byte (the last depth pass turns them into the pair
// offsets). 64-bit aligned: kWhistOffset is even.
constexpr int kWhistOffset = kVisOffset + kVisReps * kVisStride;
constexpr int kGhistWords = kWhistOffset + 2 * kHistReps * kRadix;
static_assert(kWhistOffset % 2 == 0, "64-bit sums");

__device__ __forceinline__ uint32_t depth_bits(float z) {
    uint32_t u = __float_as_uint(z);
    return (u >> 31) ? ~u : (u ^ 0x80000000u);
}

// floor(v) converted to int with the float clamped to [-1, hi] first, so an
// out-of-range extent saturates identically on every platform.
__device__ __forceinline__ int tile_floor(float v, int hi) {
    float f = floorf(v);
    f = fmaxf(f, -1.0f);
    f = fminf(f, (float)hi);
    return (int)f;
}

// Start of a forward (see ForwardStart): device scalars, a new look-back
// epoch, zero tickets, digit histograms and visible-count replicas, empty
// tile ranges, zero losses (one per instance), zero segment sums of the tile
// sort; and a fitter step's prelude.
__global__ __launch_bounds__(256) void k_begin(ForwardStart f) {
    forward_start(f, gtid(), gstride());
}

// Adds 1 to bin d of an LDS histogram for every active lane. When all active
// lanes share the bin (the upper bytes of the tile ids of consecutive pairs,
// of the depths of neighbouring splats) one lane adds their count: same-bin
// LDS atomics from a whole wave would serialise.
__device__ __forceinline__ void lds_count(uint32_t *h, uint32_t d) {
    const uint64_t active = __ballot(true);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(d == d0) == active) {
        if (__lane_id() == __ffsll((unsigned long long)active) - 1)
            atomicAdd(&h[d0], (uint32_t)__popcll(active));
    } else {
        atomicAdd(&h[d], 1u);
    }
}

// Adds w to bin d of an LDS histogram for every lane of the wave (all lanes
// active). When the lanes share the bin (the top depth byte of a run of
// splats) one lane adds the wave's sum.
__device__ __forceinline__ void lds_add_weight(uint32_t *h, uint32_t d, uint32_t w) {
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(d == d0) == __ballot(true)) {
        uint32_t sum = w;
#pragma unroll
        for (int m = 1; m < kWave; m <<= 1) sum += __shfl_xor(sum, m, kWave);
        if (__lane_id() == 0 && sum) atomicAdd(&h[d0], sum);
    } else if (w) {
        atomicAdd(&h[d], w);
    }
}

// Copy r of the per-top-byte tile-count sums in ghist.
__device__ __forceinline__ unsigned long long *whist_row(uint32_t *ghist, int r) {
    return reinterpret_cast<unsigned long long *>(ghist + kWhistOffset) + (size_t)r * kRadix;
}

// Adds a 256-thread workgroup's LDS digit histograms (rows x 256 bins; row r
// counts bits [8r, 8r + 8) of the keys) to the global rows.
__device__ __forceinline__ void flush_histograms(const uint32_t (*s_h)[kRadix], int rows,
                                                 uint32_t *grows) {
    for (int r = 0; r < rows; ++r) {
        const uint32_t c = s_h[r][threadIdx.x];
        if (c) atomicAdd(&grows[r * kRadix + threadIdx.x], c);
    }
}

// Zeroes up to ten arrays of n floats in one launch.
struct ZeroList {
    float *p[10];
};

__global__ __launch_bounds__(256) void k_zero10(ZeroList z, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
#pragma unroll
        for (int f = 0; f < 10; ++f)
            if (z.p[f]) z.p[f][i] = 0.f;
    }
}

// Reach threshold of a splat: with Q' = -(qa, qb; qb, qc) (the record's
// conic, positive definite), every point d where the compositing's float
// evaluation of the falloff can pass the alpha test has d^T Q' d <= tm, and
// the returned value bounds tm plus the rounding of the compositing's own
// test of d^T Q' d against it. A point passes when d^T Q' d <= t2 = log2(o /
// (1/255)) up to rounding: the float value of p = -d^T Q' d (two fma chains)
// is within 12 eps (tr^2 / det) |p| of the exact one (eps = 2^-24), exp2 and
// the opacity product within 2^-20 relative (1e-5 in t2); so tm = (t2 +
// 1e-5) / (1 - slack) with slack = 16 eps tr^2 / det, and the float form in
// the test is within slack of exact again. Computed in double and rounded
// up; a conic too ill-conditioned for the bound (or not positive definite in
// float) gets +inf (always evaluated), an opacity below the threshold -inf
// (never).
__device__ float reach_threshold(float qa, float qb, float qc, float o) {
    if (!(o >= kAlphaThres)) return -INFINITY;
    const double a = -(double)qa, b = -(double)qb, c = -(double)qc;
    const double det = a * c - b * b, tr = a + c;
    const double slack = 16.0 * 0x1p-24 * (tr * tr / det);
    if (!(a > 0.0 && c > 0.0 && det > 0.0 && slack < 0.25)) return INFINITY;
    const double tm = (log2((double)o / (double)kAlphaThres) + 1e-5) / (1.0 - slack);
    return (float)(tm * (1.0 + slack) * (1.0 + 1e-6) + 1e-6);
}

// Per splat: cull test and tile rectangle exactly as tileAssignKernel
// (src/layers/tile_rasterize_layer.cu:76-125), with FP contraction off so
// the decisions are reproducible on the host. With `hist` the four depth-byte
// histograms go to replica blockIdx % kHistReps of ghist (the one-kernel
// depth passes read them); the visible count goes to one of the kVisReps
// words, once per workgroup.
// With MULTI the image of splat i is that of its instance, mi[sinst[i]], and
// the rectangle is in that image's tiles.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_prep(ds_splat2d in, int n, int W, int H, int ntx,
                                              int nty, uint2 *__restrict__ rect,
                                              uint32_t *__restrict__ ntiles,
                                              uint32_t *__restrict__ dkey,
                                              uint32_t *__restrict__ didx,
                                              float4 *__restrict__ rec, uint32_t *ghist,
                                              int hist, const MultiInst *__restrict__ mi,
                                              const uint16_t *__restrict__ sinst) {
#pragma clang fp contract(off)
    __shared__ uint32_t s_h[4][kRadix];  // the four depth-byte histograms
    __shared__ uint32_t s_w[kRadix];     // tile counts per top depth byte
    __shared__ uint32_t s_vis;
    if (hist) {
#pragma unroll
        for (int r = 0; r < 4; ++r) s_h[r][threadIdx.x] = 0u;
        s_w[threadIdx.x] = 0u;
    }
    if (threadIdx.x == 0) s_vis = 0u;
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool visible = false;
    uint32_t key = 0xFFFFFFFFu;  // culled splats sort last
    uint32_t nt = 0;             // tiles of the splat
    if (i < n) {
        if constexpr (MULTI) {
            const MultiInst I = mi[sinst[i]];
            W = I.W;
            H = I.H;
            ntx = I.ntx;
            nty = I.nty;
        }
        const float x = in.pos_x[i], y = in.pos_y[i], z = in.pos_z[i];
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        float cxx = 0.f, cxy = 0.f, cyy = 0.f, det = 0.f;
        if (fabsf(x) <= 1.0f && fabsf(y) <= 1.0f && fabsf(z) <= 1.0f) {
            cxx = in.cov_xx[i];
            cxy = in.cov_xy[i];
            cyy = in.cov_yy[i];
            det = cxx * cyy - cxy * cxy;
            if (det > 0.f) {
                const float ex = 3.f * sqrtf(cxx);
                const float ey = 3.f * sqrtf(cyy);
                const float pixel_ndc = fmaxf(2.f / (float)W, 2.f / (float)H);
                if (fmaxf(ex, ey) * 2.f >= pixel_ndc) {
                    const float fx = (float)W / (float)kTile;
                    const float fy = (float)H / (float)kTile;
                    tx0 = max(tile_floor((x - ex + 1.f) * 0.5f * fx, ntx), 0);
                    tx1 = min(tile_floor((x + ex + 1.f) * 0.5f * fx, ntx), ntx - 1);
                    ty0 = max(tile_floor((1.f - (y + ey)) * 0.5f * fy, nty), 0);
                    ty1 = min(tile_floor((1.f - (y - ey)) * 0.5f * fy, nty), nty - 1);
                    visible = (tx0 <= tx1 && ty0 <= ty1);
                }
            }
        }
        didx[i] = (uint32_t)i;
        if (visible) {
            rect[i] = make_uint2((uint32_t)tx0 | ((uint32_t)tx1 << 16),
                                 (uint32_t)ty0 | ((uint32_t)ty1 << 16));
            nt = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
            ntiles[i] = nt;
            key = depth_bits(z);
            const float idet = 1.f / det;
            const float k = kFalloffScale;
            const float qa = cyy * idet * k, qb = -cxy * idet * k, qc = cxx * idet * k;
            const float o = in.opacity[i];
            rec[3 * i + 0] = make_float4(x, y, qa, qb);
            rec[3 * i + 1] = make_float4(qc, o, reach_threshold(qa, qb, qc, o), __uint_as_float(i));
            rec[3 * i + 2] = make_float4(in.color_r[i], in.color_g[i], in.color_b[i], 0.f);
        } else {
            ntiles[i] = 0;
        }
        dkey[i] = key;
        if (hist) {
#pragma unroll
            for (int r = 0; r < 4; ++r) lds_count(s_h[r], (key >> (8 * r)) & 0xFFu);
        }
    }
    if (hist) lds_add_weight(s_w, key >> 24, nt);
    const uint64_t vis = __ballot(visible);
    if (__lane_id() == 0 && vis) atomicAdd(&s_vis, (uint32_t)__popcll(vis));
    __syncthreads();
    if (threadIdx.x == 0 && s_vis)
        atomicAdd(&ghist[kVisOffset + (blockIdx.x % kVisReps) * kVisStride], s_vis);
    if (hist) {
        flush_histograms(s_h, 4, ghist + (blockIdx.x % kHistReps) * kHistRepWords);
        const uint32_t wsum = s_w[threadIdx.x];
        if (wsum)
            atomicAdd(whist_row(ghist, blockIdx.x % kHistReps) + threadIdx.x,
                      (unsigned long long)wsum);
    }
}

// Writes the pairs of the splats, taken in depth order, each splat's tiles in
// row-major order starting at its scanned offset; pairs at or past the cap
// are dropped. Slot p belongs to the last depth rank d with off[d] <= p (off
// is the exclusive scan of the tile counts in depth order, so that splat's
// range [off[d], off[d+1]) holds p and is not empty). Each wave takes a span
// of kEmitSpan consecutive slots, 64 at a time: it finds the owner of the
// span's first slot by one uniform binary search between the owners of the
// first slots of its chunk of 2^kEmitChunkShift slots and of the next chunk
// (recorded by the offset scan); every window then covers slots owned by the
// 64 depth ranks from the owner of its first slot on (each visible splat has
// a pair), so lane i reads off[lo + i], marks the slot where rank lo + i
// starts (an LDS max, per wave) and a max-scan over the lanes gives each
// slot's owner. The next window starts from the owner of the last slot. A
// splat covering thousands of tiles costs no more than thousands of small
// ones, and the dependent loads per slot are a few per window instead of a
// binary search per slot. The first thread also totals the visible-count
// replicas into the device scalar. With MULTI the tile id is global: the
// splat's instance's first tile plus the tile in its image.
constexpr int kEmitChunkShift = 12;
constexpr int kEmitSpan = 512;  // consecutive slots per wave and grid sweep

__device__ __forceinline__ uint32_t wave_incl_max_u32(uint32_t x) {
    const int l = __lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (l >= d) x = max(x, y);
    }
    return x;
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_emit(const uint32_t *__restrict__ order,
                                              const uint32_t *__restrict__ off,
                                              const uint32_t *__restrict__ chunk_first,
                                              const uint2 *__restrict__ rect, uint32_t *sc,
                                              const uint32_t *__restrict__ ghist, int ntx,
                                              uint32_t *__restrict__ tkey,
                                              uint32_t *__restrict__ tval,
                                              const MultiInst *__restrict__ mi,
                                              const uint16_t *__restrict__ sinst,
                                              uint2 *__restrict__ pair_base,
                                              uint32_t *__restrict__ slot_splat) {
    __shared__ uint32_t s_head[256 / kWave][kWave];
    const uint32_t n = sc[SC_COUNT];
    const uint32_t m = sc[SC_NISECT];  // pairs kept: the total clamped at the cap
    if (gtid() == 0) {
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < kVisReps; ++r) v += ghist[kVisOffset + r * kVisStride];
        sc[SC_VISIBLE] = v;
    }
    const int lane = __lane_id();
    uint32_t *const head = s_head[threadIdx.x / kWave];
    const uint32_t nch = (m + (1u << kEmitChunkShift) - 1) >> kEmitChunkShift;
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t s0 = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) *
                       kEmitSpan;
         s0 < m; s0 += nwaves * kEmitSpan) {  // wave-uniform
        const uint32_t s1 = (uint32_t)min((uint64_t)m, s0 + kEmitSpan);
        // the owner of the span's first slot (uniform binary search)
        const uint32_t c = (uint32_t)s0 >> kEmitChunkShift;
        uint32_t lo = chunk_first[c];
        uint32_t hi = (c + 1 < nch) ? chunk_first[c + 1] + 1 : n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= (uint32_t)s0)
                lo = mid;
            else
                hi = mid;
        }
        for (uint32_t P = (uint32_t)s0; P < s1; P += kWave) {
            // where ranks lo + 1 .. lo + 63 start, relative to the window
            const uint32_t k = lo + (uint32_t)lane;
            const int d = k < n ? (int)(off[k] - P) : kWave;
            head[lane] = 0u;
            __builtin_amdgcn_wave_barrier();
            if (lane > 0 && d >= 0 && d < kWave) atomicMax(&head[d], (uint32_t)lane);
            __builtin_amdgcn_wave_barrier();
            const uint32_t rel = wave_incl_max_u32(head[lane]);  // owner of slot P + lane
            const int d_own = __shfl(d, (int)rel, kWave);
            const uint32_t p = P + (uint32_t)lane;
            if (p < s1) {
                const uint32_t ko = lo + rel;
                const uint32_t g = order[ko];
                const uint32_t j = (uint32_t)(lane - d_own);  // index inside the splat's rectangle
                const uint2 r = rect[g];
                const uint32_t tx0 = r.x & 0xFFFF, tx1 = r.x >> 16, ty0 = r.y & 0xFFFF;
                const uint32_t w = tx1 - tx0 + 1;
                const uint32_t ty = ty0 + j / w, tx = tx0 + j % w;
                // the splat's first slot: where its backward records go (see
                // CompositeArgs)
                if (j == 0 && pair_base) pair_base[g] = make_uint2(p - (ty0 * w + tx0), w);
                if constexpr (MULTI) {
                    const MultiInst I = mi[sinst[g]];
                    tkey[p] = (uint32_t)I.tile_base + ty * (uint32_t)I.ntx + tx;
                } else {
                    tkey[p] = ty * (uint32_t)ntx + tx;
                }
                tval[p] = g;
                if (slot_splat) slot_splat[p] = g;  // kept: the tile passes reuse tval
            }
            // the next window starts with the owner of this window's last slot
            lo += (uint32_t)__builtin_amdgcn_readlane((int)rel, kWave - 1);
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Tile boundaries in the sorted pair list (buildTileRangesKernel), plus the
// optional 64-bit sorted keys.
__global__ __launch_bounds__(256) void k_ranges(const uint32_t *__restrict__ tkey,
                                                const uint32_t *__restrict__ tval,
                                                const float *__restrict__ pos_z,
                                                const uint32_t *sc, int2 *__restrict__ ranges,
                                                uint64_t *__restrict__ keys64) {
    const uint32_t n = sc[SC_NISECT];
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t t = tkey[i];
        if (i == 0 || tkey[i - 1] != t) ranges[t].x = (int)i;
        if (i == (int64_t)n - 1 || tkey[i + 1] != t) ranges[t].y = (int)(i + 1);
        if (keys64) keys64[i] = ((uint64_t)t << 32) | depth_bits(pos_z[tval[i]]);
    }
}

// ---------------------------------------------------------------- tile sort
// Stable counting sort of the depth-ordered pairs on their tile id in one
// pass, for up to kTileBinsMax tiles (instead of 8-bit radix passes). The
// pair slots are cut into chunks of kTileChunk; the pair-offset scan records
// the splat owning each chunk's first slot.
//   k_tile_emit     per chunk: emits the chunk's pairs (as k_emit), ranks
//                   each among the chunk's pairs of the same tile in slot
//                   order, writes (tile | rank << kTileIdBits) and the splat,
//                   the chunk's per-tile counts (row c of a chunk-major count
//                   matrix) and adds them to the sums of its segment (one of
//                   kTileSegs runs of consecutive chunks)
//   k_tile_segscan  per tile: exclusive scan over the segments, tile total
//   k_tile_offsets  per tile and segment: the tile's first output slot
//                   (exclusive scan of the totals), its range, and the count
//                   matrix replaced in place by each (chunk, tile)'s first slot
//   k_tile_scatter  per pair: output slot = (chunk, tile) slot + rank
// Within a tile the pairs keep their emission order, i.e. depth order.
constexpr int kTileChunk = 4096;
// 512 threads: a chunk's workgroup is alone on its CU (LDS), so its latency
// is the kernel's; more waves shorten its per-thread loops.
constexpr int kTileThreads = 512;
constexpr int kTileWaves = kTileThreads / kWave;         // 8
constexpr int kTileWaveSlots = kTileChunk / kTileWaves;  // consecutive slots per wave
constexpr int kTileGroups = kTileWaveSlots / kWave;      // 64-slot groups per wave
constexpr int kTileItems = kTileChunk / kTileThreads;    // slots per thread in the head scan
constexpr int kTileIdBits = 12;
constexpr int kTileBinsMax = 1 << kTileIdBits;
constexpr int kTileSegs = 32;

// Chunks per segment for nch chunks.
__device__ __forceinline__ uint32_t tile_seg_len(uint32_t nch) {
    return (nch + kTileSegs - 1) / kTileSegs;
}

// LDS: per-wave tile counters (64 KB at kTileBinsMax), the owner of each slot
// (16-bit: it stays below kTileChunk) and, for a chunk of at most kTileKMax
// splats, each splat's id, tile box and first slot: under 80 KB, so that two
// chunks' workgroups share a CU (one per CU left half the fit's chunks to a
// second round). A chunk of more splats (about six pairs each or fewer)
// reads a slot's owner's first slot, id and box from global memory instead.
// Memory round trips per chunk: the chunk's first and last splats; their
// first slots and ids; their boxes (two more per slot, past kTileKMax).
constexpr int kTileKMax = 640;
__global__ __launch_bounds__(kTileThreads) void k_tile_emit(
    const uint32_t *__restrict__ order, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ chunk_first, const uint2 *__restrict__ rect, uint32_t *sc,
    const uint32_t *__restrict__ ghist, int ntx, int ntiles, int tile_bits,
    uint32_t *__restrict__ pk, uint32_t *__restrict__ val, uint32_t *__restrict__ cnt,
    uint32_t *seg, uint2 *__restrict__ pair_base) {
    __shared__ uint16_t s_w[kTileWaves][kTileBinsMax];
    __shared__ uint16_t s_head[kTileChunk];
    __shared__ uint32_t s_gid[kTileKMax];
    __shared__ uint32_t s_box[kTileKMax];  // tx0 | ty0 << 10 | (tx1 - tx0) << 20
    __shared__ uint16_t s_first[kTileKMax];
    __shared__ uint32_t s_red[kTileWaves];
    __shared__ uint32_t s_first0;
    const int tid = threadIdx.x, w = tid / kWave, lane = __lane_id();
    const uint32_t n = sc[SC_COUNT];
    const uint32_t m = sc[SC_NISECT];  // pairs kept: the total clamped at the cap
    if (gtid() == 0) {
        uint32_t v = 0;
#pragma unroll
        for (int r = 0; r < kVisReps; ++r) v += ghist[kVisOffset + r * kVisStride];
        sc[SC_VISIBLE] = v;
    }
    const uint32_t nch = (m + kTileChunk - 1) / kTileChunk;
    const uint32_t G = tile_seg_len(nch);
    const uint64_t lt = lanemask_lt();
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t cs = c * kTileChunk;
        const uint32_t cc = min((uint32_t)kTileChunk, m - cs);
        for (int t = tid; t < ntiles; t += kTileThreads) {
#pragma unroll
            for (int q = 0; q < kTileWaves; ++q) s_w[q][t] = 0;
        }
        for (int i = tid; i < kTileChunk; i += kTileThreads) s_head[i] = 0;
        // The chunk's splats are the depth ranks d0..dn: the owners of its
        // first slot and of the next chunk's (the last chunk: up to n - 1).
        const uint32_t d0 = min(chunk_first[c], n - 1);
        const uint32_t dn = (c + 1 < nch) ? min(chunk_first[c + 1], n - 1) : n - 1;
        const uint32_t K = (dn >= d0) ? min((uint32_t)kTileChunk, dn - d0 + 1) : 1u;
        const bool in_lds = K <= (uint32_t)kTileKMax;  // uniform
        uint32_t o[kTileItems], g[kTileItems];  // first slots and ids of the chunk's splats
#pragma unroll
        for (int i = 0; i < kTileItems; ++i) {
            const uint32_t k = (uint32_t)(tid + i * kTileThreads);
            o[i] = g[i] = 0;
            if (k < K) {
                o[i] = off[d0 + k];
                if (in_lds) g[i] = order[d0 + k];
            }
        }
        __syncthreads();  // s_head zeroed
        // Every visible splat has at least one pair, so the splats after d0
        // start at strictly rising slots past cs; those starting at or after
        // the chunk's end (culled splats are depth-sorted last and start at
        // or after m) are not heads.
#pragma unroll
        for (int i = 0; i < kTileItems; ++i) {
            const uint32_t k = (uint32_t)(tid + i * kTileThreads);
            if (in_lds && k < K) {
                const uint2 r = rect[g[i]];
                const uint32_t tx0 = r.x & 0xFFFF, tx1 = r.x >> 16, ty0 = r.y & 0xFFFF;
                s_gid[k] = g[i];
                s_box[k] = tx0 | (ty0 << 10) | ((tx1 - tx0) << 20);
                // a splat whose first slot is in this chunk: where its
                // backward records go (see CompositeArgs)
                if (pair_base && o[i] >= cs && o[i] < cs + cc) {
                    const uint32_t wd = tx1 - tx0 + 1;
                    pair_base[g[i]] = make_uint2(o[i] - (ty0 * wd + tx0), wd);
                }
                if (k == 0)
                    s_first0 = o[i];
                else if (o[i] > cs && o[i] < cs + cc)
                    s_first[k] = (uint16_t)(o[i] - cs);
            }
            if (k > 0 && k < K && o[i] > cs && o[i] < cs + cc) s_head[o[i] - cs] = (uint16_t)k;
        }
        __syncthreads();
        // Inclusive max-scan of the heads: the owner of every slot.
        {
            uint32_t v[kTileItems];
            uint32_t mx = 0;
#pragma unroll
            for (int j = 0; j < kTileItems; ++j) {
                v[j] = s_head[tid * kTileItems + j];
                mx = max(mx, v[j]);
            }
            uint32_t incl = mx;
#pragma unroll
            for (int dd = 1; dd < kWave; dd <<= 1) {
                const uint32_t y = __shfl_up(incl, dd, kWave);
                if (lane >= dd) incl = max(incl, y);
            }
            if (lane == kWave - 1) s_red[w] = incl;
            __syncthreads();
            uint32_t run = __shfl_up(incl, 1, kWave);
            if (lane == 0) run = 0;
            for (int q = 0; q < w; ++q) run = max(run, s_red[q]);
#pragma unroll
            for (int j = 0; j < kTileItems; ++j) {
                run = max(run, v[j]);
                s_head[tid * kTileItems + j] = (uint16_t)run;
            }
        }
        __syncthreads();
        // Wave w takes slots [w * kTileWaveSlots, (w + 1) * kTileWaveSlots)
        // in groups of 64, in order.
        uint32_t tile[kTileGroups], gid[kTileGroups], rnk[kTileGroups];
        if (in_lds) {
#pragma unroll
            for (int j = 0; j < kTileGroups; ++j) {
                const uint32_t q = (uint32_t)(w * kTileWaveSlots + j * kWave + lane);
                tile[j] = gid[j] = 0;
                if (q < cc) {
                    const uint32_t k = s_head[q];
                    // the slot's index among its owner's pairs
                    const uint32_t jj = (k == 0) ? cs + q - s_first0 : q - s_first[k];
                    const uint32_t b = s_box[k];
                    const uint32_t tx0 = b & 1023u, ty0 = (b >> 10) & 1023u, wd = (b >> 20) + 1;
                    gid[j] = s_gid[k];
                    tile[j] = (ty0 + jj / wd) * (uint32_t)ntx + tx0 + jj % wd;
                }
            }
        } else {
            uint32_t jj[kTileGroups];
#pragma unroll
            for (int j = 0; j < kTileGroups; ++j) {
                const uint32_t q = (uint32_t)(w * kTileWaveSlots + j * kWave + lane);
                gid[j] = jj[j] = 0;
                if (q < cc) {
                    const uint32_t d = d0 + s_head[q];
                    gid[j] = order[d];
                    jj[j] = cs + q - off[d];
                }
            }
#pragma unroll
            for (int j = 0; j < kTileGroups; ++j) {
                const uint32_t q = (uint32_t)(w * kTileWaveSlots + j * kWave + lane);
                tile[j] = 0;
                if (q < cc) {
                    const uint2 r = rect[gid[j]];
                    const uint32_t tx0 = r.x & 0xFFFF, tx1 = r.x >> 16, ty0 = r.y & 0xFFFF;
                    const uint32_t wd = tx1 - tx0 + 1;
                    tile[j] = (ty0 + jj[j] / wd) * (uint32_t)ntx + tx0 + jj[j] % wd;
                    // the owner's first slot: where its backward records go
                    if (pair_base && jj[j] == 0)
                        pair_base[gid[j]] = make_uint2(cs + q - (ty0 * wd + tx0), wd);
                }
            }
        }
        // Rank among the wave's earlier slots of the same tile: lanes of one
        // tile found by matching the tile bits with ballots; the group's
        // first such lane advances the wave's counter.
#pragma unroll
        for (int j = 0; j < kTileGroups; ++j) {
            const uint32_t q = (uint32_t)(w * kTileWaveSlots + j * kWave + lane);
            const bool valid = q < cc;
            const uint32_t t = tile[j];
            uint64_t peers = __ballot(valid);
            for (int bit = 0; bit < tile_bits; ++bit) {
                const bool on = (t >> bit) & 1u;
                const uint64_t bb = __ballot(on);
                peers &= on ? bb : ~bb;
            }
            const uint32_t r = (uint32_t)__popcll(peers & lt);
            uint32_t old = 0;
            if (valid) old = s_w[w][t];
            if (valid && r == 0) s_w[w][t] = (uint16_t)(old + (uint32_t)__popcll(peers));
            rnk[j] = old + r;
        }
        __syncthreads();
        // Per tile: each wave's first rank, the chunk's count, its segment sum.
        for (int t = tid; t < ntiles; t += kTileThreads) {
            uint32_t run = 0;
#pragma unroll
            for (int q = 0; q < kTileWaves; ++q) {
                const uint32_t v = s_w[q][t];
                s_w[q][t] = (uint16_t)run;
                run += v;
            }
            cnt[(size_t)c * ntiles + t] = run;
            if (run) atomicAdd(&seg[(size_t)(c / G) * ntiles + t], run);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kTileGroups; ++j) {
            const uint32_t q = (uint32_t)(w * kTileWaveSlots + j * kWave + lane);
            if (q < cc) {
                const uint32_t t = tile[j];
                pk[cs + q] = t | ((s_w[w][t] + rnk[j]) << kTileIdBits);
                val[cs + q] = gid[j];
            }
        }
        __syncthreads();  // the LDS is reused by the next chunk
    }
}

// Per tile: segment sums -> their exclusive scan (in place), and the total.
__global__ __launch_bounds__(256) void k_tile_segscan(uint32_t *__restrict__ seg,
                                                      uint32_t *__restrict__ tot, int ntiles) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t v[kTileSegs];
#pragma unroll
    for (int s = 0; s < kTileSegs; ++s) v[s] = seg[(size_t)s * ntiles + t];
    uint32_t run = 0;
#pragma unroll
    for (int s = 0; s < kTileSegs; ++s) {
        seg[(size_t)s * ntiles + t] = run;
        run += v[s];
    }
    tot[t] = run;
}

// Workgroup (x, s): tiles [256 x, 256 x + 256) in segment s. The tile's first
// output slot is the exclusive scan of the totals; segment 0 writes the
// ranges of the non-empty tiles (empty ones stay (0, 0)) and, with inv, every
// tile's range at its position in the dispatch order; the counts of the
// segment's chunks become first slots.
__global__ __launch_bounds__(256) void k_tile_offsets(uint32_t *__restrict__ cnt,
                                                      const uint32_t *__restrict__ seg,
                                                      const uint32_t *__restrict__ tot,
                                                      const uint32_t *sc, int ntiles,
                                                      int2 *__restrict__ ranges,
                                                      const uint32_t *__restrict__ inv,
                                                      uint32_t *__restrict__ order,
                                                      int2 *__restrict__ order_range) {
    __shared__ uint32_t s_tmp[256 / kWave];
    const uint32_t m = sc[SC_NISECT];
    const uint32_t nch = (m + kTileChunk - 1) / kTileChunk;
    const uint32_t G = tile_seg_len(nch);
    const uint32_t s = blockIdx.y;
    const int t0 = blockIdx.x * 256;
    const int t = t0 + threadIdx.x;
    uint32_t before = 0;
    for (int i = threadIdx.x; i < t0; i += 256) before += tot[i];
    uint32_t base;
    block_excl_scan<256>(before, s_tmp, &base);
    const uint32_t mine = (t < ntiles) ? tot[t] : 0u;
    const uint32_t start = base + block_excl_scan<256>(mine, s_tmp, nullptr);
    if (t >= ntiles) return;
    const int2 rt = mine ? make_int2((int)start, (int)(start + mine)) : make_int2(0, 0);
    if (s == 0 && mine) ranges[t] = rt;
    // a dispatch order computed before these ranges (NextOrder): each tile
    // and its range at its position
    if (s == 0 && inv) {
        const uint32_t pos = inv[t];
        order[pos] = (uint32_t)t;
        order_range[pos] = rt;
    }
    uint32_t run = start + seg[(size_t)s * ntiles + t];
    const uint32_t c1 = min(s * G + G, nch);
    for (uint32_t c = s * G; c < c1; c += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (c + j < c1) ? cnt[(size_t)(c + j) * ntiles + t] : 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (c + j < c1) {
                cnt[(size_t)(c + j) * ntiles + t] = run;
                run += v[j];
            }
        }
    }
}

// Per pair: its sorted position; optionally the 64-bit (tile | depth) key.
__global__ __launch_bounds__(256) void k_tile_scatter(const uint32_t *__restrict__ pk,
                                                      const uint32_t *__restrict__ val,
                                                      const uint32_t *__restrict__ cnt,
                                                      const uint32_t *sc, int ntiles,
                                                      const float *__restrict__ pos_z,
                                                      uint32_t *__restrict__ flat,
                                                      uint64_t *__restrict__ keys64) {
    const uint32_t m = sc[SC_NISECT];
    for (int64_t p = gtid(); p < m; p += gstride()) {
        const uint32_t x = pk[p];
        const uint32_t t = x & (uint32_t)(kTileBinsMax - 1);
        const uint32_t c = (uint32_t)p / kTileChunk;
        if (t >= (uint32_t)ntiles) continue;  // never, for consistent counts
        const uint32_t dst = cnt[(size_t)c * ntiles + t] + (x >> kTileIdBits);
        const uint32_t g = val[p];
        if (dst < m) {  // always, for consistent counts
            flat[dst] = g;
            if (keys64) keys64[dst] = ((uint64_t)t << 32) | depth_bits(pos_z[g]);
        }
    }
}

// Dispatch order of the compositing workgroups, heaviest tiles first: tiles
// bucketed by floor(log2(pairs)) + 1 (empty tiles: bucket 0, 2^30 and more
// pairs share bucket 31), buckets in descending order (the order inside a
// bucket is arbitrary; results do not depend on it). One workgroup. Pass 1
// reads the ranges (kOrderUnroll loads in flight per thread), keeps each
// tile's bucket in LDS when the tiles fit (kOrderLdsTiles) and counts the
// buckets per wave: the lanes of one bucket are found with five ballots and
// their first lane adds their number to the wave's row (distinct buckets,
// distinct words: no atomics). Pass 2 reserves each wave's slots with one
// LDS atomic per bucket, and writes the order, its inverse and each
// position's range. (A fitter's steps get the same order from the pair
// reduction of the step before instead: see ds_raster::lag_order.)
constexpr int kOrderThreads = 1024;
constexpr int kOrderWaves = kOrderThreads / kWave;
constexpr int kOrderUnroll = 4;
constexpr int kOrderLdsTiles = 49152;
__global__ __launch_bounds__(kOrderThreads) void k_tile_order(const int2 *__restrict__ ranges,
                                                              int ntiles,
                                                              uint32_t *__restrict__ order,
                                                              uint32_t *__restrict__ inv,
                                                              int2 *__restrict__ order_range) {
    __shared__ uint32_t h[32];
    __shared__ uint32_t wh[kOrderWaves][32];
    __shared__ uint8_t s_b[kOrderLdsTiles];
    const int tid = threadIdx.x, w = tid / kWave;
    const bool cache = ntiles <= kOrderLdsTiles;
    const uint64_t lt = lanemask_lt();
    for (int i = tid; i < kOrderWaves * 32; i += kOrderThreads) (&wh[0][0])[i] = 0u;
    __syncthreads();
    constexpr int kStep = kOrderThreads * kOrderUnroll;
    for (int t0 = 0; t0 < ntiles; t0 += kStep) {
        int2 r[kOrderUnroll];
#pragma unroll
        for (int u = 0; u < kOrderUnroll; ++u) {
            const int t = t0 + u * kOrderThreads + tid;
            r[u] = t < ntiles ? ranges[t] : make_int2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < kOrderUnroll; ++u) {
            const int t = t0 + u * kOrderThreads + tid;
            const bool valid = t < ntiles;
            const uint32_t b = tile_bucket(r[u]);
            if (valid && cache) s_b[t] = (uint8_t)b;
            const uint64_t peers = same_bucket(b, valid);
            if (valid && (peers & lt) == 0) wh[w][b] += (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
    if (tid < 32) {  // bucket totals
        uint32_t v = 0;
        for (int q = 0; q < kOrderWaves; ++q) v += wh[q][tid];
        h[tid] = v;
    }
    __syncthreads();
    if (tid == 0) {  // descending exclusive scan: the first slot of each bucket
        uint32_t run = 0;
        for (int b = 31; b >= 0; --b) {
            const uint32_t v = h[b];
            h[b] = run;
            run += v;
        }
    }
    __syncthreads();
    for (int t0 = 0; t0 < ntiles; t0 += kOrderThreads) {
        const int t = t0 + tid;
        const bool valid = t < ntiles;
        const int2 rt = valid ? ranges[t] : make_int2(0, 0);
        const uint32_t b = !valid ? 0u : cache ? (uint32_t)s_b[t] : tile_bucket(rt);
        const uint64_t peers = same_bucket(b, valid);
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        uint32_t base = 0;
        if (valid && rank == 0) base = atomicAdd(&h[b], (uint32_t)__popcll(peers));
        const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : 0;
        base = (uint32_t)__shfl((int)base, leader, kWave);
        if (valid) {
            order[base + rank] = (uint32_t)t;
            inv[t] = base + rank;
            order_range[base + rank] = rt;  // read with the order: one load level less
        }
    }
}

int grid_cap(uint64_t n, int per, int cap) {
    uint64_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return b > (uint64_t)cap ? cap : (int)b;
}

}  // namespace
}  // namespace ds

// ====================================================================== object

struct ds_raster {
    int W = 0, H = 0, ntx = 0, nty = 0, ntiles = 0;
    int count_cap = 0;
    uint32_t isect_cap = 0;
    int keep_keys = 0;
    int grad_mode = DS_GRAD_EXACT;
    int fwd_ppt = 1;  // compositing forward: pixels per thread (1, 2 or 4)
    int bwd_ppt = 1;  // compositing backward: pixels per thread
    int tile_order = DS_TILE_ORDER_HEAVY_FIRST;  // compositing dispatch order
    int depth_sort = DS_DEPTH_SORT_ONESWEEP;     // how the splats are sorted by depth
    float depth_sort_ms[3] = {-1.f, -1.f, -1.f};  // the last autotune's times per mode
    int tile_bits = 1;
    int tile_passes = 1;  // 8-bit radix passes over the tile ids
    // Several fit instances in one rasterizer (ninst > 0): splats, tiles and
    // pixels of instance k are contiguous runs of the global index spaces;
    // per instance its layout, per splat and per tile its instance.
    int ninst = 0;
    int64_t npix = 0;  // pixels of all instances
    ds::DevBuf<ds::MultiInst> mtab;
    ds::DevBuf<uint16_t> sinst, tinst;
    // per splat
    ds::DevBuf<uint2> rect;
    ds::DevBuf<uint32_t> ntiles_of, dkey_a, dkey_b, didx_a, didx_b, pair_off;
    ds::DevBuf<float4> rec;
    // per pair
    ds::DevBuf<uint32_t> tkey_a, tkey_b, tval_a, tval_b;
    ds::DevBuf<uint64_t> keys64;
    // backward: per pair slot its gradient record (3 float4), per splat its
    // record base and tile-rectangle width (see CompositeArgs); released for
    // forward-only use
    ds::DevBuf<float4> pair_rec;
    ds::DevBuf<uint2> pair_base;
    ds::DevBuf<uint32_t> pair_splat;  // radix tile sort: each slot's splat
    // one-pass tile sort (up to kTileBinsMax tiles): chunk x tile counts,
    // segment x tile sums, tile totals, the splat owning each chunk's first slot
    ds::DevBuf<uint32_t> tcnt, tseg, ttot, tfirst;
    // per tile / pixel
    ds::DevBuf<int2> ranges;
    ds::DevBuf<uint32_t> order, order_inv;  // compositing dispatch order (heavy-first), inverse
    ds::DevBuf<int2> order_range;  // the range of tile order[b], at b
    ds::DevBuf<float> T;
    ds::DevBuf<int32_t> last;
    // scratch: device scalars, depth-digit histograms, look-back status words
    // of the depth passes and of the pair-offset scan, and the per-chunk
    // histograms and scan partials of the tile passes
    ds::DevBuf<uint32_t> scalars, ghist, hist, part;
    ds::DevBuf<uint64_t> lb_depth;
    // results of the last forward
    bool order_valid = false;  // `order` was computed by the last forward
    // The fitter step's dispatch order: the pair reduction computes each
    // tile's position (order_inv) from the ranges of the step before
    // (NextOrder), and the next forward's k_tile_offsets writes the order and
    // the positions' ranges from it, instead of k_tile_order (one launch less
    // per step; the pair counts change little between steps).
    bool lag_order = false;
    const uint32_t *flat = nullptr;
    // each pair slot's splat before the tile sort (slots of a splat are
    // contiguous): the emitted values, or pair_splat where the tile sort
    // overwrites them
    const uint32_t *slot_splat = nullptr;
    bool base_valid = false;  // the last forward wrote pair_base
    int last_count = -1;
    bool has_fwd = false;
};

namespace ds {

// The one-pass tile sort handles up to kTileBinsMax tiles of one instance.
static bool one_pass_tile_sort(const ds_raster *r) {
    return r->ninst == 0 && r->ntiles <= kTileBinsMax;
}

// The dispatch order comes from the step before (see ds_raster::lag_order).
static bool lagged_order(const ds_raster *r) {
    static_assert(kTileBinsMax <= kNextOrderMaxTiles, "one-pass tile counts fit NextOrder");
    return r->lag_order && one_pass_tile_sort(r) && r->tile_order == DS_TILE_ORDER_HEAVY_FIRST;
}

static ds_status raster_alloc_tiles(ds_raster *r);

static ds_status raster_alloc_pixels(ds_raster *r, int W, int H) {
    DS_REQUIRE(W > 0 && H > 0 && W <= 16384 && H <= 16384, "raster size out of range");
    r->W = W;
    r->H = H;
    r->ntx = (W + kTile - 1) / kTile;
    r->nty = (H + kTile - 1) / kTile;
    r->ntiles = r->ntx * r->nty;
    r->npix = (int64_t)W * H;
    return raster_alloc_tiles(r);
}

// Tile-sort and per-tile / per-pixel buffers for r->ntiles tiles and r->npix
// pixels.
static ds_status raster_alloc_tiles(ds_raster *r) {
    int bits = 1;
    while ((1 << bits) < r->ntiles) ++bits;
    r->tile_bits = bits;
    r->tile_passes = (bits + kRadixBits - 1) / kRadixBits;
    const size_t max_chunks = ((size_t)r->isect_cap + kTileChunk - 1) / kTileChunk;
    DS_TRY(r->tfirst.alloc(max_chunks));
    if (one_pass_tile_sort(r)) {
        DS_TRY(r->tcnt.alloc(max_chunks * (size_t)r->ntiles));
        DS_TRY(r->tseg.alloc((size_t)kTileSegs * r->ntiles));
        DS_TRY(r->ttot.alloc((size_t)r->ntiles));
    } else {
        r->tcnt.release();
        r->tseg.release();
        r->ttot.release();
    }
    DS_TRY(r->ranges.alloc((size_t)r->ntiles));
    DS_TRY(r->order.alloc((size_t)r->ntiles));
    DS_TRY(r->order_inv.alloc((size_t)r->ntiles));
    DS_TRY(r->order_range.alloc((size_t)r->ntiles));
    {  // the dispatch order is a permutation from the start: the identity
        std::vector<uint32_t> id((size_t)r->ntiles);
        for (size_t t = 0; t < id.size(); ++t) id[t] = (uint32_t)t;
        if (!id.empty()) {
            DS_TRY_HIP(hipMemcpy(r->order.ptr, id.data(), id.size() * 4, hipMemcpyHostToDevice));
            DS_TRY_HIP(
                hipMemcpy(r->order_inv.ptr, id.data(), id.size() * 4, hipMemcpyHostToDevice));
        }
    }
    DS_TRY(r->T.alloc((size_t)r->npix));
    DS_TRY(r->last.alloc((size_t)r->npix));
    r->has_fwd = false;
    return DS_OK;
}

static CompositeArgs composite_args(const ds_raster *r) {
    const bool multi = r->ninst > 0;
    return CompositeArgs{(const float4 *)r->rec.ptr, r->flat, (const int2 *)r->ranges.ptr,
                         r->W, r->H, r->ntx, r->ntiles,
                         multi ? (const MultiInst *)r->mtab.ptr : nullptr,
                         multi ? (const uint16_t *)r->tinst.ptr : nullptr,
                         r->tile_order == DS_TILE_ORDER_HEAVY_FIRST ? (const uint32_t *)r->order.ptr
                                                                    : nullptr,
                         r->pair_rec.ptr, (const uint2 *)r->pair_base.ptr,
                         (const uint32_t *)r->scalars.ptr + SC_EPOCH,
                         r->tile_order == DS_TILE_ORDER_HEAVY_FIRST
                             ? (const int2 *)r->order_range.ptr
                             : nullptr};
}

// The backward needs the gradient records and the pair bases of the last
// forward.
static ds_status require_backward(const ds_raster *r) {
    DS_REQUIRE(r->pair_rec.ptr && r->pair_base.ptr,
               "the rasterizer's backward buffers were released (ds_raster_release_backward)");
    DS_REQUIRE(r->base_valid || r->last_count == 0,
               "the last forward ran without the backward buffers");
    return DS_OK;
}

// Each splat's records of this epoch into grad (after a backward compositing);
// grad_is_zero when the same forward's first kernel cleared it.
static ds_status reduce_pairs(ds_raster *r, const ds_splat2d &grad, hipStream_t s,
                              bool grad_is_zero = false) {
    if (r->last_count <= 0 || !r->slot_splat) return DS_OK;
    uint32_t *sc = r->scalars.ptr;
    const NextOrder next = lagged_order(r) ? NextOrder{(const int2 *)r->ranges.ptr, r->ntiles,
                                                       r->order_inv.ptr}
                                           : NextOrder{nullptr, 0, nullptr};
    return launch_pair_reduce(r->pair_rec.ptr, r->slot_splat, sc + SC_NISECT, sc + SC_EPOCH,
                              (int64_t)r->isect_cap, grad, grad_is_zero, s, next);
}

// Tile assignment, sorts, ranges and dispatch order, then compositing: the
// forward alone, or (fused_grad non-NULL, target given, equal launch shapes)
// the forward and backward of each tile in one kernel accumulating into
// *fused_grad.
// The first kernel's work for n splats: the forward's scalars and counters,
// and (grad_zero, tick non-NULL) a fitter step's prelude.
static ForwardStart forward_start_of(ds_raster *r, int n, float *loss,
                                     const ds_splat2d *grad_zero, int32_t *tick) {
    uint32_t *sc = r->scalars.ptr;
    const bool multi = r->ninst > 0;
    ForwardStart f{};
    f.zero[0] = sc;  // SC_NISECT .. SC_EPOCH - 1, with SC_COUNT = n
    f.nzero[0] = SC_EPOCH;
    f.set_word = sc + SC_COUNT;
    f.set_value = (uint32_t)n;
    f.zero[1] = sc + SC_TICKET;  // the tickets (not the sticky error flag)
    f.nzero[1] = LB_SLOTS;
    f.zero[2] = loss ? (uint32_t *)loss : nullptr;
    f.nzero[2] = loss ? (uint32_t)(multi ? r->ninst : 1) : 0u;
    f.zero[3] = r->ghist.ptr;
    f.nzero[3] = kGhistWords;
    f.zero[4] = (uint32_t *)r->ranges.ptr;
    f.nzero[4] = 2u * (uint32_t)r->ntiles;
    f.zero[5] = r->tseg.ptr;
    f.nzero[5] = one_pass_tile_sort(r) ? (uint32_t)(kTileSegs * r->ntiles) : 0u;
    f.epoch = sc + SC_EPOCH;
    f.tick = tick;
    if (grad_zero) {
        const ds_splat2d &g = *grad_zero;
        float *arrs[10] = {g.pos_x, g.pos_y,   g.pos_z,   g.cov_xx,  g.cov_xy,
                           g.cov_yy, g.color_r, g.color_g, g.color_b, g.opacity};
        for (int k = 0; k < 10; ++k) f.grad[k] = arrs[k];
        f.ngrad = g.count;
    }
    return f;
}

static ds_status raster_forward(ds_raster *r, const ds_splat2d &in, float *out_rgb,
                                const float *target, float *loss, float *grad, hipStream_t s,
                                const ds_splat2d *fused_grad = nullptr,
                                const ds_splat2d *grad_zero = nullptr, int32_t *tick = nullptr,
                                bool started = false, bool sort_only = false) {
    const int n = in.count;
    DS_REQUIRE(n >= 0 && n <= r->count_cap, "splat count exceeds the rasterizer's count_cap");
    DS_REQUIRE(r->ninst == 0 || n == r->count_cap,
               "a multi-instance rasterizer takes exactly the splats of all its instances");
    DS_REQUIRE(sort_only || out_rgb != nullptr, "out_rgb is NULL");
    DS_REQUIRE((target == nullptr) == (loss == nullptr) && (target == nullptr) == (grad == nullptr),
               "target, loss and grad must be given together");
    uint32_t *sc = r->scalars.ptr;
    uint32_t *gh = r->ghist.ptr;
    // look-back slot `slot` with its status words at `status`
    auto slot = [&](uint64_t *status, int lb) {
        return LookbackSlot{status, sc + SC_TICKET + lb, sc + SC_EPOCH, sc + SC_ERR};
    };
    const bool one_pass_tiles = one_pass_tile_sort(r);
    const bool multi = r->ninst > 0;
    const MultiInst *mi = multi ? r->mtab.ptr : nullptr;
    const uint16_t *si = multi ? r->sinst.ptr : nullptr;
    if (!started) {
        const ForwardStart f = forward_start_of(r, n, loss, grad_zero, tick);
        uint64_t words = (uint64_t)f.ngrad;
        for (int k = 0; k < kStartRanges; ++k) words = std::max<uint64_t>(words, f.nzero[k]);
        hipLaunchKernelGGL(k_begin, dim3(grid_cap(words / 4, 256, 1024)), dim3(256), 0, s, f);
        DS_CHECK_LAUNCH("begin");
    }
    if (n > 0) {
        const bool one_kernel =
            r->depth_sort == DS_DEPTH_SORT_ONESWEEP && (uint32_t)n <= kOneKernelMaxKeys;
        const int dbits = (!one_kernel && r->depth_sort == DS_DEPTH_SORT_RADIX11)
                              ? kWideRadixBits
                              : kRadixBits;
        const int dpasses = (32 + dbits - 1) / dbits;
        if (multi)
            hipLaunchKernelGGL(k_prep<true>, dim3((n + 255) / 256), dim3(256), 0, s, in, n, r->W,
                               r->H, r->ntx, r->nty, r->rect.ptr, r->ntiles_of.ptr,
                               r->dkey_a.ptr, r->didx_a.ptr, r->rec.ptr, gh, one_kernel ? 1 : 0,
                               mi, si);
        else
            hipLaunchKernelGGL(k_prep<false>, dim3((n + 255) / 256), dim3(256), 0, s, in, n,
                               r->W, r->H, r->ntx, r->nty, r->rect.ptr, r->ntiles_of.ptr,
                               r->dkey_a.ptr, r->didx_a.ptr, r->rec.ptr, gh, one_kernel ? 1 : 0,
                               mi, si);
        DS_CHECK_LAUNCH("prep");
        // Depth sort of splats: LSD passes over the 32 depth bits (four of 8
        // bits, or three of 11), and the pair offsets in depth order (total
        // and capped count). Up to kOneKernelMaxKeys splats each 8-bit pass
        // can be one kernel with look-back (launches dominate there), the
        // last pass also writing the offsets; otherwise reduce-then-scan
        // passes and a separate scan (the look-back traffic grows with the
        // chunk count). The autotuners time the choices.
        uint32_t *ka = r->dkey_a.ptr, *kb = r->dkey_b.ptr, *va = r->didx_a.ptr,
                 *vb = r->didx_b.ptr;
        const size_t dstride = lookback_words((uint64_t)r->count_cap, kRadix);
        const uint64_t max_chunks = ((uint64_t)r->isect_cap + kTileChunk - 1) / kTileChunk;
        static_assert(kTileChunk == 1 << 12, "chunk heads use shift 12");
        static_assert(kEmitChunkShift == 12, "one chunk size for both emissions");
        const ScanHeads heads{r->tfirst.ptr, 12u, (uint32_t)max_chunks};
        for (int pass = 0; pass < dpasses; ++pass) {
            if (one_kernel) {
                // the last pass also writes the pair offsets in depth order
                const PairOffsets po =
                    pass == 3 ? PairOffsets{r->ntiles_of.ptr,
                                            (const unsigned long long *)(gh + kWhistOffset),
                                            kHistReps,
                                            r->pair_off.ptr,
                                            r->lb_depth.ptr + 4 * dstride,
                                            sc + SC_WANTED,
                                            sc + SC_NISECT,
                                            r->isect_cap,
                                            heads}
                              : PairOffsets{nullptr, nullptr, 0, nullptr, nullptr,
                                            nullptr, nullptr, 0, ScanHeads{nullptr, 0, 0}};
                DS_TRY(onesweep_pass(ka, va, kb, vb, sc + SC_COUNT, (uint32_t)n, 8 * pass,
                                     gh + (LB_DEPTH + pass) * kRadix, kHistReps,
                                     kHistRepWords,
                                     slot(r->lb_depth.ptr + pass * dstride, LB_DEPTH + pass), s,
                                     po));
            } else {
                DS_TRY(radix_pass<uint32_t>(ka, va, kb, vb, sc + SC_COUNT, (uint32_t)n,
                                            dbits * pass, r->hist.ptr, r->part.ptr, s, dbits));
            }
            std::swap(ka, kb);
            std::swap(va, vb);
        }
        const uint32_t *order = va;
        if (!one_kernel)
            DS_TRY(scan_exclusive_u32(r->ntiles_of.ptr, order, r->pair_off.ptr,
                                      ScanLen{sc + SC_COUNT, 1, 1}, (uint32_t)n, r->part.ptr,
                                      sc + SC_WANTED, sc + SC_NISECT, r->isect_cap, s, heads));
        if (one_pass_tiles) {
            hipLaunchKernelGGL(k_tile_emit, dim3(grid_cap(max_chunks, 1, 1024)),
                               dim3(kTileThreads), 0, s, order,
                               (const uint32_t *)r->pair_off.ptr,
                               (const uint32_t *)r->tfirst.ptr, (const uint2 *)r->rect.ptr, sc,
                               (const uint32_t *)gh, r->ntx, r->ntiles, r->tile_bits,
                               r->tkey_a.ptr, r->tval_a.ptr, r->tcnt.ptr, r->tseg.ptr,
                               r->pair_base.ptr);
            DS_CHECK_LAUNCH("tile_emit");
            const int tblocks = (r->ntiles + 255) / 256;
            hipLaunchKernelGGL(k_tile_segscan, dim3(tblocks), dim3(256), 0, s, r->tseg.ptr,
                               r->ttot.ptr, r->ntiles);
            DS_CHECK_LAUNCH("tile_segscan");
            const bool lag = lagged_order(r);
            hipLaunchKernelGGL(k_tile_offsets, dim3(tblocks, kTileSegs), dim3(256), 0, s,
                               r->tcnt.ptr, (const uint32_t *)r->tseg.ptr,
                               (const uint32_t *)r->ttot.ptr, (const uint32_t *)sc, r->ntiles,
                               r->ranges.ptr, lag ? (const uint32_t *)r->order_inv.ptr : nullptr,
                               lag ? r->order.ptr : (uint32_t *)nullptr,
                               lag ? r->order_range.ptr : (int2 *)nullptr);
            DS_CHECK_LAUNCH("tile_offsets");
            hipLaunchKernelGGL(k_tile_scatter, dim3(grid_cap(r->isect_cap, 256, 4096)),
                               dim3(256), 0, s, (const uint32_t *)r->tkey_a.ptr,
                               (const uint32_t *)r->tval_a.ptr, (const uint32_t *)r->tcnt.ptr,
                               (const uint32_t *)sc, r->ntiles, (const float *)in.pos_z,
                               r->tval_b.ptr, r->keep_keys ? r->keys64.ptr : (uint64_t *)nullptr);
            DS_CHECK_LAUNCH("tile_scatter");
            r->flat = r->tval_b.ptr;
            r->slot_splat = r->tval_a.ptr;  // the emitted values stay
        } else {
            const dim3 egrid(grid_cap((uint64_t)r->isect_cap, 256, 4096));
            if (multi)
                hipLaunchKernelGGL(k_emit<true>, egrid, dim3(256), 0, s, order,
                                   (const uint32_t *)r->pair_off.ptr,
                                   (const uint32_t *)r->tfirst.ptr, (const uint2 *)r->rect.ptr,
                                   sc, (const uint32_t *)gh, r->ntx, r->tkey_a.ptr,
                                   r->tval_a.ptr, mi, si, r->pair_base.ptr, r->pair_splat.ptr);
            else
                hipLaunchKernelGGL(k_emit<false>, egrid, dim3(256), 0, s, order,
                                   (const uint32_t *)r->pair_off.ptr,
                                   (const uint32_t *)r->tfirst.ptr, (const uint2 *)r->rect.ptr,
                                   sc, (const uint32_t *)gh, r->ntx, r->tkey_a.ptr,
                                   r->tval_a.ptr, mi, si, r->pair_base.ptr, r->pair_splat.ptr);
            DS_CHECK_LAUNCH("emit");
            r->slot_splat = r->pair_splat.ptr;
            // Stable sort of the pairs on the tile bits. With hundreds of
            // chunks the reduce-then-scan passes beat one-kernel passes,
            // whose look-back traffic grows with the chunk count.
            uint32_t *tka = r->tkey_a.ptr, *tkb = r->tkey_b.ptr, *tva = r->tval_a.ptr,
                     *tvb = r->tval_b.ptr;
            for (int pass = 0; pass < r->tile_passes; ++pass) {
                DS_TRY(radix_pass<uint32_t>(tka, tva, tkb, tvb, sc + SC_NISECT, r->isect_cap,
                                            kRadixBits * pass, r->hist.ptr, r->part.ptr, s));
                std::swap(tka, tkb);
                std::swap(tva, tvb);
            }
            r->flat = tva;
            hipLaunchKernelGGL(k_ranges, dim3(grid_cap(r->isect_cap, 256, 4096)), dim3(256), 0,
                               s, (const uint32_t *)tka, (const uint32_t *)tva,
                               (const float *)in.pos_z, (const uint32_t *)sc, r->ranges.ptr,
                               r->keep_keys ? r->keys64.ptr : (uint64_t *)nullptr);
            DS_CHECK_LAUNCH("ranges");
        }
    } else {
        r->flat = r->tval_a.ptr;
    }
    if (r->tile_order == DS_TILE_ORDER_HEAVY_FIRST && !lagged_order(r)) {
        hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(kOrderThreads), 0, s,
                           (const int2 *)r->ranges.ptr, r->ntiles, r->order.ptr,
                           r->order_inv.ptr, r->order_range.ptr);
        DS_CHECK_LAUNCH("tile_order");
    }
    r->order_valid = r->tile_order == DS_TILE_ORDER_HEAVY_FIRST;
    r->last_count = n;
    r->base_valid = r->pair_base.ptr != nullptr;
    if (sort_only) return DS_OK;
    r->has_fwd = true;
    const CompositeArgs ca = composite_args(r);
    if (fused_grad) DS_TRY(require_backward(r));
    if (fused_grad && target && r->fwd_ppt == r->bwd_ppt)
        DS_TRY(launch_composite_fused(ca, r->fwd_ppt, r->grad_mode, out_rgb, r->T.ptr,
                                      r->last.ptr, target, loss, grad, s));
    else
        DS_TRY(launch_composite_fwd(ca, r->fwd_ppt, out_rgb, r->T.ptr, r->last.ptr, target,
                                    loss, grad, s));
    if (fused_grad && !(target && r->fwd_ppt == r->bwd_ppt))
        DS_TRY(launch_composite_bwd(ca, r->bwd_ppt, r->grad_mode, r->T.ptr, r->last.ptr,
                                    grad, s));
    // a fitter step's prelude (grad_zero) cleared this gradient
    if (fused_grad)
        DS_TRY(reduce_pairs(r, *fused_grad, s,
                            grad_zero != nullptr && grad_zero->pos_x == fused_grad->pos_x));
    return DS_OK;
}

ds_status raster_step_fwd_bwd(ds_raster *r, const ds_splat2d &in, float *out_rgb,
                              const float *target, float *loss, float *grad_rgb,
                              const ds_splat2d &grad_in, int32_t *adam_tick, hipStream_t s,
                              bool started) {
    DS_REQUIRE(r, "NULL raster");
    DS_REQUIRE(target && loss && grad_rgb, "the fitter step needs the MSE target");
    DS_REQUIRE(grad_in.count == in.count, "grad_in count differs from the input's");
    r->lag_order = true;  // each step's reduction orders the next step's tiles
    return raster_forward(r, in, out_rgb, target, loss, grad_rgb, s, &grad_in, &grad_in,
                          adam_tick, started);
}

ds_status raster_step_start(ds_raster *r, int n, float *loss, const ds_splat2d &grad_in,
                            int32_t *adam_tick, ForwardStart *out) {
    DS_REQUIRE(r && out, "NULL argument");
    DS_REQUIRE(n >= 0 && n <= r->count_cap, "splat count exceeds the rasterizer's count_cap");
    *out = forward_start_of(r, n, loss, &grad_in, adam_tick);
    return DS_OK;
}

static ds_status raster_backward(ds_raster *r, const ds_splat2d &in, const float *grad_rgb,
                                 const ds_splat2d &gin, hipStream_t s) {
    DS_REQUIRE(r->has_fwd, "ds_raster_bwd called before ds_raster_fwd");
    DS_REQUIRE(in.count == r->last_count, "splat count differs from the last forward");
    DS_REQUIRE(grad_rgb != nullptr, "grad_rgb is NULL");
    DS_TRY(require_backward(r));
    CompositeArgs ca = composite_args(r);
    if (!r->order_valid) ca.order = nullptr, ca.order_range = nullptr;  // order switched on after the forward
    DS_TRY(launch_composite_bwd(ca, r->bwd_ppt, r->grad_mode, r->T.ptr, r->last.ptr,
                                grad_rgb, s));
    return reduce_pairs(r, gin, s);
}

}  // namespace ds

// ====================================================================== C-ABI

using namespace ds;

// Per-splat and per-pair buffers, scratch and device scalars of a raster
// whose pixel/tile buffers are allocated; frees r on failure.
static ds_status raster_finish_create(ds_raster *r, ds_status st, ds_raster **out) {
    const size_t nc = (size_t)(r->count_cap > 0 ? r->count_cap : 1);
    const int64_t cap = r->isect_cap;
    if (st == DS_OK) st = r->rect.alloc(nc);
    if (st == DS_OK) st = r->ntiles_of.alloc(nc);
    if (st == DS_OK) st = r->dkey_a.alloc(nc);
    if (st == DS_OK) st = r->dkey_b.alloc(nc);
    if (st == DS_OK) st = r->didx_a.alloc(nc);
    if (st == DS_OK) st = r->didx_b.alloc(nc);
    if (st == DS_OK) st = r->pair_off.alloc(nc);
    if (st == DS_OK) st = r->rec.alloc(3 * nc);
    if (st == DS_OK) st = r->tkey_a.alloc(cap);
    if (st == DS_OK) st = r->tkey_b.alloc(cap);
    if (st == DS_OK) st = r->tval_a.alloc(cap);
    if (st == DS_OK) st = r->tval_b.alloc(cap);
    if (st == DS_OK && r->keep_keys) st = r->keys64.alloc(cap);
    // backward records start with tag 0: epoch 0 is never used
    if (st == DS_OK) st = r->pair_rec.alloc(3 * (size_t)cap);
    if (st == DS_OK) st = fill_u32((uint32_t *)r->pair_rec.ptr, 0u, 12 * (size_t)cap, nullptr);
    if (st == DS_OK) st = r->pair_base.alloc(nc);
    if (st == DS_OK) st = r->pair_splat.alloc(cap);
    // look-back status words start at zero: epoch 0 is never used
    // four depth passes and the last pass's pair-offset sums
    const size_t lbd = 5 * lookback_words(nc, kRadix);
    if (st == DS_OK) st = r->lb_depth.alloc(lbd);
    if (st == DS_OK) st = fill_u32((uint32_t *)r->lb_depth.ptr, 0u, 2 * lbd, nullptr);
    if (st == DS_OK) st = r->ghist.alloc((size_t)kGhistWords);
    if (st == DS_OK)
        st = r->hist.alloc(std::max<uint64_t>(
            sort_hist_words(std::max<uint64_t>((uint64_t)cap, (uint64_t)nc)),
            sort_hist_words((uint64_t)nc, kWideRadixBits)));
    if (st == DS_OK) st = r->part.alloc(kSpineMax);
    if (st == DS_OK) st = r->scalars.alloc(SC_NUM);
    if (st == DS_OK) st = fill_u32(r->scalars.ptr, 0u, SC_NUM, nullptr);
    if (st == DS_OK) st = fill_u32(r->ranges.ptr, 0u, (size_t)r->ntiles * 2, nullptr);
    if (st == DS_OK && hipDeviceSynchronize() != hipSuccess)
        st = set_error(DS_ERR_HIP, "hipDeviceSynchronize after raster create");
    if (st != DS_OK) {
        delete r;
        return st;
    }
    *out = r;
    return DS_OK;
}

extern "C" ds_status ds_raster_create(const ds_raster_config *cfg, ds_raster **out) {
    clear_error();
    DS_REQUIRE(cfg && out, "NULL argument");
    DS_REQUIRE(cfg->count_cap >= 0, "count_cap < 0");
    DS_REQUIRE((uint64_t)cfg->count_cap <= kLookbackMaxCount, "count_cap above 2^30 - 1");
    int64_t cap = cfg->isect_cap > 0 ? cfg->isect_cap : (int64_t)1 << 25;
    DS_REQUIRE(cap <= (int64_t)kLookbackMaxCount, "isect_cap above 2^30 - 1");
    DS_REQUIRE(cfg->grad_mode == DS_GRAD_EXACT || cfg->grad_mode == DS_GRAD_REFERENCE,
               "grad_mode must be DS_GRAD_EXACT or DS_GRAD_REFERENCE");
    auto *r = new ds_raster();
    r->count_cap = cfg->count_cap;
    r->isect_cap = (uint32_t)cap;
    r->depth_sort = (uint32_t)r->count_cap <= kOneKernelMaxKeys ? DS_DEPTH_SORT_ONESWEEP
                                                                 : DS_DEPTH_SORT_RADIX8;
    r->keep_keys = cfg->keep_keys;
    r->grad_mode = cfg->grad_mode;
    return raster_finish_create(r, raster_alloc_pixels(r, cfg->width, cfg->height), out);
}

extern "C" ds_status ds_raster_create_multi(const ds_raster_instance *inst, int n_inst,
                                            int64_t isect_cap, int keep_keys, int grad_mode,
                                            ds_raster **out) {
    clear_error();
    DS_REQUIRE(inst && out, "NULL argument");
    DS_REQUIRE(n_inst >= 1 && n_inst <= 65535, "n_inst must be in [1, 65535]");
    DS_REQUIRE(grad_mode == DS_GRAD_EXACT || grad_mode == DS_GRAD_REFERENCE,
               "grad_mode must be DS_GRAD_EXACT or DS_GRAD_REFERENCE");
    int64_t cap = isect_cap > 0 ? isect_cap : (int64_t)1 << 25;
    DS_REQUIRE(cap <= (int64_t)kLookbackMaxCount, "isect_cap above 2^30 - 1");
    std::vector<MultiInst> tab((size_t)n_inst);
    int64_t nsplat = 0, ntile = 0, npix = 0;
    for (int k = 0; k < n_inst; ++k) {
        const ds_raster_instance &q = inst[k];
        DS_REQUIRE(q.width > 0 && q.height > 0 && q.width <= 16384 && q.height <= 16384,
                   "instance size out of range");
        DS_REQUIRE(q.count >= 0, "instance splat count < 0");
        MultiInst &m = tab[(size_t)k];
        m.W = q.width;
        m.H = q.height;
        m.ntx = (q.width + kTile - 1) / kTile;
        m.nty = (q.height + kTile - 1) / kTile;
        m.tile_base = (int)ntile;
        m.pix_base = (int)npix;
        m.inv_npix = 1.f / ((float)q.width * (float)q.height);
        m.pad = 0;
        nsplat += q.count;
        ntile += (int64_t)m.ntx * m.nty;
        npix += (int64_t)q.width * q.height;
        DS_REQUIRE(nsplat <= (int64_t)kLookbackMaxCount, "total splat count above 2^30 - 1");
        DS_REQUIRE(ntile < ((int64_t)1 << 30) && npix < ((int64_t)1 << 31) / 3,
                   "instances too large for one rasterizer");
    }
    std::vector<uint16_t> sinst((size_t)std::max<int64_t>(nsplat, 1), 0),
        tinst((size_t)ntile, 0);
    for (int k = 0, s0 = 0; k < n_inst; s0 += inst[k].count, ++k) {
        std::fill(sinst.begin() + s0, sinst.begin() + s0 + inst[k].count, (uint16_t)k);
        const MultiInst &m = tab[(size_t)k];
        std::fill(tinst.begin() + m.tile_base, tinst.begin() + m.tile_base + m.ntx * m.nty,
                  (uint16_t)k);
    }
    auto *r = new ds_raster();
    r->count_cap = (int)nsplat;
    r->isect_cap = (uint32_t)cap;
    r->depth_sort = (uint32_t)r->count_cap <= kOneKernelMaxKeys ? DS_DEPTH_SORT_ONESWEEP
                                                                 : DS_DEPTH_SORT_RADIX8;
    r->keep_keys = keep_keys;
    r->grad_mode = grad_mode;
    r->ninst = n_inst;
    r->ntiles = (int)ntile;
    r->npix = npix;
    ds_status st = raster_alloc_tiles(r);
    if (st == DS_OK) st = r->mtab.alloc(tab.size());
    if (st == DS_OK) st = r->sinst.alloc(sinst.size());
    if (st == DS_OK) st = r->tinst.alloc(tinst.size());
    if (st == DS_OK &&
        (hipMemcpy(r->mtab.ptr, tab.data(), tab.size() * sizeof(MultiInst),
                   hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(r->sinst.ptr, sinst.data(), sinst.size() * 2, hipMemcpyHostToDevice) !=
             hipSuccess ||
         hipMemcpy(r->tinst.ptr, tinst.data(), tinst.size() * 2, hipMemcpyHostToDevice) !=
             hipSuccess))
        st = set_error(DS_ERR_HIP, "upload of the instance tables");
    return raster_finish_create(r, st, out);
}

extern "C" ds_status ds_raster_splat_instances(ds_raster *r, const uint16_t **dev) {
    clear_error();
    DS_REQUIRE(r && dev, "NULL argument");
    DS_REQUIRE(r->ninst > 0, "not a multi-instance rasterizer");
    *dev = r->sinst.ptr;
    return DS_OK;
}

extern "C" ds_status ds_raster_destroy(ds_raster *r) {
    clear_error();
    delete r;
    return DS_OK;
}

extern "C" ds_status ds_raster_resize(ds_raster *r, int width, int height) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    DS_REQUIRE(r->ninst == 0, "a multi-instance rasterizer cannot be resized");
    if (width == r->W && height == r->H) return DS_OK;
    return raster_alloc_pixels(r, width, height);
}

extern "C" ds_status ds_raster_fwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                                   const float *target_rgb, float *loss_dev, float *grad_rgb,
                                   ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in, "NULL argument");
    return raster_forward(r, *in, out_rgb, target_rgb, loss_dev, grad_rgb, as_stream(s));
}

extern "C" ds_status ds_raster_fwd_bwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                                       const float *target_rgb, float *loss_dev,
                                       float *grad_rgb, const ds_splat2d *grad_in,
                                       ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in && grad_in, "NULL argument");
    DS_REQUIRE(target_rgb && loss_dev && grad_rgb, "ds_raster_fwd_bwd needs the MSE target");
    DS_REQUIRE(grad_in->count == in->count, "grad_in count differs from the input's");
    return raster_forward(r, *in, out_rgb, target_rgb, loss_dev, grad_rgb, as_stream(s),
                          grad_in);
}

extern "C" ds_status ds_raster_composite_fwd_bwd(ds_raster *r, const ds_splat2d *in,
                                                 float *out_rgb, const float *target_rgb,
                                                 float *loss_dev, float *grad_rgb,
                                                 const ds_splat2d *grad_in, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in && grad_in, "NULL argument");
    DS_REQUIRE(out_rgb && target_rgb && loss_dev && grad_rgb, "NULL image argument");
    DS_REQUIRE(r->has_fwd, "ds_raster_composite_fwd_bwd called before a forward");
    DS_REQUIRE(in->count == r->last_count, "splat count differs from the last forward");
    DS_REQUIRE(r->fwd_ppt == r->bwd_ppt,
               "the fused compositing needs equal forward and backward launch shapes");
    DS_TRY(require_backward(r));
    CompositeArgs ca = composite_args(r);
    if (!r->order_valid) ca.order = nullptr, ca.order_range = nullptr;
    hipStream_t st = as_stream(s);
    DS_TRY(fill_u32(loss_dev, 0u, r->ninst > 0 ? (size_t)r->ninst : 1, st));
    DS_TRY(launch_composite_fused(ca, r->fwd_ppt, r->grad_mode, out_rgb, r->T.ptr, r->last.ptr,
                                  target_rgb, loss_dev, grad_rgb, st));
    return reduce_pairs(r, *grad_in, st);
}

extern "C" ds_status ds_raster_release_backward(ds_raster *r) {
    clear_error();
    DS_REQUIRE(r, "NULL argument");
    DS_TRY_HIP(hipDeviceSynchronize());
    r->pair_rec.release();
    r->pair_base.release();
    r->pair_splat.release();
    r->base_valid = false;
    return DS_OK;
}

extern "C" ds_status ds_raster_composite_fwd(ds_raster *r, const ds_splat2d *in, float *out_rgb,
                                             ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in && out_rgb, "NULL argument");
    DS_REQUIRE(r->has_fwd, "ds_raster_composite_fwd called before a forward");
    DS_REQUIRE(in->count == r->last_count, "splat count differs from the last forward");
    CompositeArgs ca = composite_args(r);
    if (!r->order_valid) ca.order = nullptr, ca.order_range = nullptr;
    return launch_composite_fwd(ca, r->fwd_ppt, out_rgb, r->T.ptr, r->last.ptr, nullptr, nullptr,
                                nullptr, as_stream(s));
}

extern "C" ds_status ds_raster_bwd(ds_raster *r, const ds_splat2d *in, const float *grad_rgb,
                                   const ds_splat2d *grad_in, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in && grad_in, "NULL argument");
    return raster_backward(r, *in, grad_rgb, *grad_in, as_stream(s));
}

extern "C" ds_status ds_raster_counts(ds_raster *r, ds_stream_t s, uint32_t *n_isects,
                                      uint32_t *visible_count, uint32_t *n_wanted) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    uint32_t h[SC_NUM];
    DS_TRY_HIP(hipMemcpyAsync(h, r->scalars.ptr, sizeof(h), hipMemcpyDeviceToHost, as_stream(s)));
    DS_TRY_HIP(hipStreamSynchronize(as_stream(s)));
    if (n_isects) *n_isects = h[SC_NISECT];
    if (visible_count) *visible_count = h[SC_VISIBLE];
    if (n_wanted) *n_wanted = h[SC_WANTED];
    if (h[SC_ERR]) {
        DS_TRY_HIP(hipMemsetAsync(r->scalars.ptr + SC_ERR, 0, sizeof(uint32_t), as_stream(s)));
        DS_TRY_HIP(hipStreamSynchronize(as_stream(s)));
        return set_error(DS_ERR_STATE, kLookbackError);
    }
    return DS_OK;
}

namespace ds {
const uint32_t *raster_splat_tiles(ds_raster *r) { return r ? r->ntiles_of.ptr : nullptr; }

ds_status raster_take_error(ds_raster *r, hipStream_t s) {
    if (!r) return DS_OK;
    uint32_t err = 0;
    DS_TRY_HIP(hipMemcpyAsync(&err, r->scalars.ptr + SC_ERR, sizeof(err), hipMemcpyDeviceToHost,
                              s));
    DS_TRY_HIP(hipStreamSynchronize(s));
    if (!err) return DS_OK;
    DS_TRY_HIP(hipMemsetAsync(r->scalars.ptr + SC_ERR, 0, sizeof(uint32_t), s));
    DS_TRY_HIP(hipStreamSynchronize(s));
    return set_error(DS_ERR_STATE, kLookbackError);
}
}  // namespace ds

extern "C" ds_status ds_raster_get_view(ds_raster *r, ds_raster_view *v) {
    clear_error();
    DS_REQUIRE(r && v, "NULL argument");
    v->isect_keys_sorted = r->keep_keys ? r->keys64.ptr : nullptr;
    v->flatten_ids = r->flat;
    v->tile_ranges = reinterpret_cast<const int32_t *>(r->ranges.ptr);
    v->transmittance = r->T.ptr;
    v->last_ids = r->last.ptr;
    v->n_isects = r->scalars.ptr + SC_NISECT;
    v->n_isects_wanted = r->scalars.ptr + SC_WANTED;
    v->visible_count = r->scalars.ptr + SC_VISIBLE;
    v->lookback_error = r->scalars.ptr + SC_ERR;
    v->num_tiles_x = r->ntx;
    v->num_tiles_y = r->nty;
    v->width = r->W;
    v->height = r->H;
    v->tile_order = r->order_valid ? r->order.ptr : nullptr;
    return DS_OK;
}

extern "C" ds_status ds_raster_set_tuning(ds_raster *r, int fwd_ppt, int bwd_ppt) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    auto ok = [](int p) { return p == 1 || p == 2 || p == 4; };
    DS_REQUIRE(ok(fwd_ppt) && ok(bwd_ppt), "pixels per thread must be 1, 2 or 4");
    r->fwd_ppt = fwd_ppt;
    r->bwd_ppt = bwd_ppt;
    return DS_OK;
}

extern "C" ds_status ds_raster_set_tile_order(ds_raster *r, int order) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    DS_REQUIRE(order == DS_TILE_ORDER_ROWS || order == DS_TILE_ORDER_HEAVY_FIRST,
               "order must be DS_TILE_ORDER_ROWS or DS_TILE_ORDER_HEAVY_FIRST");
    r->tile_order = order;
    return DS_OK;
}

extern "C" ds_status ds_raster_set_depth_sort(ds_raster *r, int mode) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    DS_REQUIRE(mode == DS_DEPTH_SORT_ONESWEEP || mode == DS_DEPTH_SORT_RADIX8 ||
                   mode == DS_DEPTH_SORT_RADIX11,
               "mode must be DS_DEPTH_SORT_ONESWEEP, _RADIX8 or _RADIX11");
    DS_REQUIRE(mode != DS_DEPTH_SORT_ONESWEEP || (uint32_t)r->count_cap <= kOneKernelMaxKeys,
               "DS_DEPTH_SORT_ONESWEEP sorts at most 131072 splats");
    DS_REQUIRE(mode != DS_DEPTH_SORT_RADIX11 ||
                   (uint64_t)sort_hist_words((uint64_t)r->count_cap, kWideRadixBits) <=
                       (uint64_t)kSpineMax * kScanChunk,
               "DS_DEPTH_SORT_RADIX11 sorts at most 2^24 splats");
    r->depth_sort = mode;
    return DS_OK;
}

extern "C" ds_status ds_raster_get_depth_sort(ds_raster *r, int *mode) {
    clear_error();
    DS_REQUIRE(r && mode, "NULL argument");
    *mode = r->depth_sort;
    return DS_OK;
}

extern "C" ds_status ds_raster_sort(ds_raster *r, const ds_splat2d *in, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(r && in, "NULL argument");
    return raster_forward(r, *in, nullptr, nullptr, nullptr, nullptr, as_stream(s), nullptr,
                          nullptr, nullptr, false, true);
}

extern "C" ds_status ds_raster_autotune_depth_sort(ds_raster *r, const ds_splat2d *in, int reps,
                                                   ds_stream_t st, int *mode, float *ms) {
    clear_error();
    DS_REQUIRE(r && in, "NULL argument");
    if (reps < 1) reps = 1;
    hipStream_t s = as_stream(st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DS_TRY_HIP(hipEventCreate(&e0));
    DS_TRY_HIP(hipEventCreate(&e1));
    const int cur = r->depth_sort;
    float t[3] = {-1.f, -1.f, -1.f};
    ds_status st_out = DS_OK;
    for (int m = 0; m < 3 && st_out == DS_OK; ++m) {
        if (ds_raster_set_depth_sort(r, m) != DS_OK) continue;  // not for this raster
        st_out = raster_forward(r, *in, nullptr, nullptr, nullptr, nullptr, s, nullptr, nullptr,
                                nullptr, false, true);  // warm-up
        if (st_out != DS_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) st_out = set_error(DS_ERR_HIP, "hipEventRecord");
        for (int k = 0; k < reps && st_out == DS_OK; ++k)
            st_out = raster_forward(r, *in, nullptr, nullptr, nullptr, nullptr, s, nullptr,
                                    nullptr, nullptr, false, true);
        if (st_out == DS_OK && (hipEventRecord(e1, s) != hipSuccess ||
                                hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&t[m], e0, e1) != hipSuccess))
            st_out = set_error(DS_ERR_HIP, "timing the depth sort");
        t[m] /= (float)reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st_out != DS_OK) {
        r->depth_sort = cur;
        return st_out;
    }
    clear_error();
    int best = cur;  // kept unless another mode is at least 3% faster
    for (int m = 0; m < 3; ++m)
        if (t[m] >= 0.f && t[m] < 0.97f * t[best]) best = m;
    r->depth_sort = best;
    for (int m = 0; m < 3; ++m) r->depth_sort_ms[m] = t[m];
    if (mode) *mode = best;
    if (ms)
        for (int m = 0; m < 3; ++m) ms[m] = t[m];
    return DS_OK;
}

extern "C" ds_status ds_raster_depth_sort_times(ds_raster *r, float *ms) {
    clear_error();
    DS_REQUIRE(r && ms, "NULL argument");
    for (int m = 0; m < 3; ++m) ms[m] = r->depth_sort_ms[m];
    return DS_OK;
}

extern "C" ds_status ds_raster_get_tile_order(ds_raster *r, int *order) {
    clear_error();
    DS_REQUIRE(r && order, "NULL argument");
    *order = r->tile_order;
    return DS_OK;
}

extern "C" ds_status ds_raster_get_tuning(ds_raster *r, int *fwd_ppt, int *bwd_ppt) {
    clear_error();
    DS_REQUIRE(r, "NULL raster");
    if (fwd_ppt) *fwd_ppt = r->fwd_ppt;
    if (bwd_ppt) *bwd_ppt = r->bwd_ppt;
    return DS_OK;
}

extern "C" ds_status ds_splat2d_zero(const ds_splat2d *v, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(v, "NULL view");
    DS_REQUIRE(v->count >= 0, "negative count");
    if (v->count == 0) return DS_OK;
    ZeroList z{{v->pos_x, v->pos_y, v->pos_z, v->cov_xx, v->cov_xy, v->cov_yy, v->color_r,
                v->color_g, v->color_b, v->opacity}};
    hipLaunchKernelGGL(k_zero10, dim3(grid_cap((uint64_t)v->count, 256, 2048)), dim3(256), 0,
                       as_stream(s), z, v->count);
    DS_CHECK_LAUNCH("zero10");
    return DS_OK;
}

extern "C" ds_status ds_sort_pairs_u64(const uint64_t *keys_in, const uint32_t *vals_in,
                                       uint64_t *keys_out, uint32_t *vals_out, int64_t n,
                                       int key_bits, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(n >= 0 && n <= ((int64_t)1 << 30), "n out of range");
    DS_REQUIRE(key_bits >= 1 && key_bits <= 64, "key_bits out of range");
    if (n == 0) return DS_OK;
    DS_REQUIRE(keys_in && vals_in && keys_out && vals_out, "NULL argument");
    hipStream_t st = as_stream(s);
    DevBuf<uint64_t> kb;
    DevBuf<uint32_t> vb, hist, part, nd;
    DS_TRY(kb.alloc((size_t)n));
    DS_TRY(vb.alloc((size_t)n));
    DS_TRY(hist.alloc(sort_hist_words((uint64_t)n)));
    DS_TRY(part.alloc(kSpineMax));
    DS_TRY(nd.alloc(1));
    DS_TRY(sort_pairs_u64_scratch(keys_in, vals_in, keys_out, vals_out, (uint32_t)n, key_bits,
                                  SortPairsScratch{kb.ptr, vb.ptr, hist.ptr, part.ptr, nd.ptr},
                                  st));
    DS_TRY_HIP(hipStreamSynchronize(st));
    return DS_OK;
}
