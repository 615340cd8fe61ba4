// Front-to-back alpha compositing per 16x16 tile and its analytic backward.
//
// Reference semantics (src/layers/tile_rasterize_layer.cu):
//   compositeFwdKernel  192-309: pixel centre in NDC, alpha = min(0.99, o e^-sigma),
//                                skip alpha < 1/255, stop once T <= 1e-4 (after
//                                accumulating), last_id = absolute sorted index
//   compositeBwdKernel  356-578: back to front from each pixel's last_id,
//                                T_before = T / max(1 - alpha, 1e-6), conic
//                                gradients via the reformulated chain rule
//
// MI355X layout: a workgroup owns one tile and each thread P vertically
// adjacent pixels of it, so a wave covers a contiguous 16 x 4P band (P = 4:
// one wave64 per tile; P = 2: two waves; P = 1: four waves). Splat
// records are staged in LDS in batches of one record per thread and read as
// broadcasts. The per-pixel work has no branches: a skipped splat is a
// select that adds exact zeros. In the backward the nine per-splat gradient
// sums are reduced across the wave by a transposed DPP reduction that leaves
// one field per lane (no LDS traffic), amortised over P pixels per lane; the
// tile's waves add theirs in LDS, and each reached (tile, splat) pair's sums
// go to one gradient record, which k_pair_reduce adds per splat (no global
// float atomics in the compositing).
// The transmittance is stored directly (not as 1 - T) so the backward's
// division chain starts from an exact value.
#include "composite.hpp"

namespace ds {
namespace {

// Tile of this workgroup: the dispatch order puts the tiles with the most
// pairs first (the longest jobs start first and the grid ends evenly), or
// is row-major.
__device__ __forceinline__ int tile_of_block(const CompositeArgs &a) {
    return a.order ? (int)a.order[blockIdx.x] : (int)blockIdx.x;
}

// Its range in the sorted list (read beside the order, not through it).
__device__ __forceinline__ int2 range_of_block(const CompositeArgs &a) {
    return a.order ? a.order_range[blockIdx.x] : a.ranges[blockIdx.x];
}

// u = k (Sigma^-1 d)_x, w = k (Sigma^-1 d)_y for d = splat centre - pixel
// and the record's scale k, so that p = dx u + dy w = k dist2,
// d dist2/d(x,y) = 2 (u, w) / k, and the covariance gradients of the
// reference (-1/4 ddx^2, -1/4 ddy^2, -2 b dist2 - 2 dx dy / det) are
// -u^2, -w^2 and -2 u w, each over k^2 (the last by u w = k^2 (b dist2 +
// dx dy / det)).
struct AlphaEval {
    float u, w, p, g, og;
    // alpha = min(0.99, o g); alpha >= 1/255 exactly when o g >= 1/255. Only
    // called where eval_alpha passed, so og is a number in [1/255, inf): the
    // median of (og, 0, 0.99) is that minimum in one instruction (fminf
    // quiets NaNs first, one more).
    __device__ __forceinline__ float alpha() const {
        return __builtin_amdgcn_fmed3f(og, 0.f, kMaxAlpha);
    }
};

// Gaussian falloff of one splat at one pixel, without branches. Forward and
// backward run this same code so their discrete decisions agree bit for bit.
// Returns false when the splat is skipped (sigma < 0 or alpha < 1/255,
// tested as o g < 1/255 since 0.99 > 1/255); g and og are then meaningless
// (g may be inf) and callers select them away rather than multiplying by
// zero.
__device__ __forceinline__ void eval_falloff(const float4 &A, const float4 &B, float px,
                                             float py, AlphaEval &e) {
    const float dx = A.x - px;
    const float dy = A.y - py;
    e.u = fmaf(A.z, dx, A.w * dy);
    e.w = fmaf(A.w, dx, B.x * dy);
    e.p = fmaf(dx, e.u, dy * e.w);  // k dist2 = -sigma / ln 2
    e.g = __builtin_amdgcn_exp2f(e.p);
    e.og = B.y * e.g;
}

__device__ __forceinline__ bool eval_alpha(const float4 &A, const float4 &B, float px, float py,
                                           AlphaEval &e) {
    eval_falloff(A, B, px, py, e);
    return (e.p <= 0.f) & (e.og >= kAlphaThres);  // sigma >= 0
}

// The wave mask of eval_alpha's test (bit l: lane l's pixel passes), taken as
// two ballots so the compiler keeps both comparisons in scalar masks.
__device__ __forceinline__ uint64_t alpha_mask(const AlphaEval &e) {
    return __builtin_amdgcn_ballot_w64(e.p <= 0.f) &
           __builtin_amdgcn_ballot_w64(e.og >= kAlphaThres);
}

__device__ __forceinline__ float pixel_ndc_x(int j, int W) {
    return 2.f * ((float)j + 0.5f) / (float)W - 1.f;
}
__device__ __forceinline__ float pixel_ndc_y(int i, int H) {
    return 1.f - 2.f * ((float)i + 0.5f) / (float)H;
}

// DPP-permuted copy of v, all lanes enabled; a lane whose source lies outside
// the pattern reads 0 (bound_ctrl), so `v + dpp(v)` folds into one
// v_add_f32_dpp.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// One halving step of a transposed reduction: lanes whose selector bit is
// clear keep field a, the others field b, each summed with the DPP partner
// (which has the opposite selector bit).
template <int CTRL>
__device__ __forceinline__ float pair_step(float a, float b, bool hi) {
    const float sa = a + dpp<CTRL>(a);
    const float sb = b + dpp<CTRL>(b);
    return hi ? sb : sa;
}

// Sums over each 16-lane row of a wave of nine per-lane values v[0..8], all
// lanes active. Four DPP steps inside each row (partners lane^1, lane^2,
// lane-4 and lane-8 in the row) each halve the fields a lane holds
// (9 -> 5 -> 3 -> 2 -> 1). Result: lanes 0, 8, 4, 2, 10, 1, 9, 5, 3 of every
// row hold their row's sums of fields 0..8 (field_of_row_lane); the other lanes
// hold garbage.
__device__ __forceinline__ float row_sum9_transposed(const float (&v)[9], int lane) {
    const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4, b3 = lane & 8;
    constexpr int kXor1 = 0xB1, kXor2 = 0x4E, kRor4 = 0x124, kRor8 = 0x128;
    const float r0 = pair_step<kXor1>(v[0], v[5], b0);
    const float r1 = pair_step<kXor1>(v[1], v[6], b0);
    const float r2 = pair_step<kXor1>(v[2], v[7], b0);
    const float r3 = pair_step<kXor1>(v[3], v[8], b0);
    const float r4 = v[4] + dpp<kXor1>(v[4]);  // field 4 in even lanes
    const float s0 = pair_step<kXor2>(r0, r3, b1);
    const float s1 = pair_step<kXor2>(r1, r4, b1);
    const float s2 = r2 + dpp<kXor2>(r2);      // valid where b1 == 0
    const float t0 = pair_step<kRor4>(s0, s2, b2);
    const float t1 = s1 + dpp<kRor4>(s1);      // valid where b2 == 0
    return pair_step<kRor8>(t0, t1, b3);
}

// The same over all 64 lanes: the row sums, then two permlane swaps add the
// four rows, so every row holds the wave's sums.
__device__ __forceinline__ float wave_sum9_transposed(const float (&v)[9], int lane) {
    float z = row_sum9_transposed(v, lane);
    // rows 0+1 and 2+3, then the two halves
    const auto q = __builtin_amdgcn_permlane16_swap(__float_as_uint(z), __float_as_uint(z), false,
                                                    false);
    z = __uint_as_float(q[0]) + __uint_as_float(q[1]);
    const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(z), __float_as_uint(z), false,
                                                    false);
    return __uint_as_float(h[0]) + __uint_as_float(h[1]);
}

// Field whose row sum a lane holds after row_sum9_transposed (-1: none); row
// 0's lanes hold the same fields of the wave sum after wave_sum9_transposed.
__device__ __forceinline__ int field_of_row_lane(int lane) {
    switch (lane & 15) {
        case 0: return 0;
        case 8: return 1;
        case 4: return 2;
        case 2: return 3;
        case 10: return 4;
        case 1: return 5;
        case 9: return 6;
        case 5: return 7;
        case 3: return 8;
        default: return -1;
    }
}

// The factor that turns a field's sum into the gradient: the sums of the
// position and covariance fields are taken with m = -2 dL/d dist2 and the
// record's scaled u, w (k = kFalloffScale), so position x, y get -1/k
// (d dist2/d(x, y) = 2 (u, w) / k), xx and yy get 1/(2 k^2) and the
// symmetric xy 1/k^2; colour and opacity sums are the gradients.
__device__ __forceinline__ float field_scale(int f) {
    constexpr float kInvK = 1.f / kFalloffScale;
    return (f == 0 || f == 1)   ? -kInvK
           : (f == 2 || f == 4) ? 0.5f * kInvK * kInvK
           : f == 3             ? kInvK * kInvK
                                : 1.f;
}

// The image a tile belongs to: its size, tiles per row, the tile's index in
// it, where its pixels start in the (concatenated) pixel arrays, and which
// instance it is (0 for one instance).
struct TileImage {
    int W, H, ntx, local, pix_base, inst;
    float inv_npix;
};

template <bool MULTI>
__device__ __forceinline__ TileImage tile_image(const CompositeArgs &a, int tile,
                                                float inv_npix) {
    if constexpr (MULTI) {
        const int k = a.tinst[tile];
        const MultiInst I = a.mi[k];
        return TileImage{I.W, I.H, I.ntx, tile - I.tile_base, I.pix_base, k, I.inv_npix};
    } else {
        return TileImage{a.W, a.H, a.ntx, tile, 0, 0, inv_npix};
    }
}

// Pixels of one thread: column j (NDC px), P rows (NDC py), inside the image.
// Wave w owns rows [4 P w, 4 P (w + 1)) of the tile, as P bands of 16 x 4
// pixels: pixel k of lane l lies in band k, row 4 P w + 4 k + l / 16. The
// compositing loops test, skip and record reach per band.
template <int P>
struct TilePixels {
    int j;
    float px;
    int row[P];
    float py[P];
    bool inside[P];
};

template <int P>
__device__ __forceinline__ TilePixels<P> tile_pixels(const TileImage &im, int tid) {
    TilePixels<P> q;
    const int trow = im.local / im.ntx, tcol = im.local % im.ntx;
    q.j = tcol * kTile + (tid & 15);
    q.px = pixel_ndc_x(q.j, im.W);
    const int w = tid / kWave, lrow = (tid & (kWave - 1)) >> 4;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        q.row[k] = trow * kTile + w * 4 * P + 4 * k + lrow;
        q.py[k] = pixel_ndc_y(q.row[k], im.H);
        q.inside[k] = q.row[k] < im.H && q.j < im.W;
    }
    return q;
}

// Whether a splat's reach ellipse (record: centre A.xy, conic A.z, A.w, B.x
// scaled so that p = -q(d), threshold B.z) meets a band of pixel centres
// [wx0, wx1] x [wy0, wy1]: the least q(d) = a dx^2 + 2 b dx dy + c dy^2 over
// the band (d = centre - pixel) is at most B.z. q is convex, so the least
// value is 0 when the centre is in the band and otherwise the least of the
// four edges' minima (each edge's 1-D minimiser clamped to the edge). The
// threshold covers the rounding of this float evaluation. A conic that is
// not positive definite has threshold +inf; any NaN on the way counts as a
// hit.
__device__ __forceinline__ float edge_min(float a, float b, float c, float rc, float dx, float lo,
                                          float hi) {
    const float dy = fminf(fmaxf(-b * dx * rc, lo), hi);
    return fmaf(dx, fmaf(2.f * b, dy, a * dx), c * dy * dy);
}

__device__ __forceinline__ bool reaches_band(const float4 &A, const float4 &B, float wx0,
                                             float wx1, float wy0, float wy1) {
    const float a = -A.z, b = -A.w, c = -B.x;
    const float dxl = A.x - wx1, dxh = A.x - wx0, dyl = A.y - wy1, dyh = A.y - wy0;
    const float ra = __builtin_amdgcn_rcpf(a), rc = __builtin_amdgcn_rcpf(c);
    // edges dx = dxl, dxh (minimising over dy), then dy = dyl, dyh (over dx)
    const float mx =
        fminf(edge_min(a, b, c, rc, dxl, dyl, dyh), edge_min(a, b, c, rc, dxh, dyl, dyh));
    const float my =
        fminf(edge_min(c, b, a, ra, dyl, dxl, dxh), edge_min(c, b, a, ra, dyh, dxl, dxh));
    float m = fminf(mx, my);
    const bool inside = (dxl <= 0.f) & (dxh >= 0.f) & (dyl <= 0.f) & (dyh >= 0.f);
    m = inside ? 0.f : m;
    return !(m > B.z);
}

// Per band (16 x 4 pixels of a wave), one bit per splat of the tile's list
// for its first 64 * kSkipWords splats: did the forward find a pixel of the
// band that the splat reaches (alpha >= 1/255, pixel not yet finished)? The
// backward's test for the same (band, splat) is the same test (a pixel's
// last contributor is where it finished), so it skips the cleared bits
// without evaluating them, and evaluates a splat only on the bands it
// reaches.
constexpr int kSkipWords = 64;

// Stages an entry of the tile's list in LDS slot t: the splat's record and,
// with SLOT, in the colour part's spare word the pair slot of (splat, this
// tile) where the backward writes the pair's gradient record (see
// CompositeArgs), so that no global load sits between a reached pair's
// sums and their store.
// g is the entry's splat (flat[idx]), read a batch ahead.
template <bool SLOT>
__device__ __forceinline__ void stage_record(const CompositeArgs &a, uint32_t g, int t,
                                             uint32_t trow, uint32_t tcol, float4 *sA,
                                             float4 *sB, float4 *sC) {
    const float4 *r = a.rec + 3 * (size_t)g;
    float4 c = r[2];
    if constexpr (SLOT) {
        const uint2 pb = a.pair_base[g];
        c.w = __uint_as_float(pb.x + trow * pb.y + tcol);
    }
    sA[t] = r[0];
    sB[t] = r[1];
    sC[t] = c;
}

// Front to back over the tile's list (compositeFwdKernel). With BITS the
// per-band reach bits are recorded in s_bits. Every exit is wave-uniform (a
// ballot), so the loop runs on scalar branches; lanes whose pixels are done
// ride along adding exact zeros. A wave whose pixels are all done skips the
// batch, and leaves it once the last of them finishes: only a splat that
// reaches a pixel can finish it.
// Returns the first list index of the batch left in sA/sB/sC (-1 if none).
template <int P, bool BITS>
__device__ __forceinline__ int composite_front_to_back(
    const CompositeArgs &a, int2 range, const TilePixels<P> &q, float4 *sA, float4 *sB,
    float4 *sC, uint64_t (*s_bits)[P][kSkipWords], float (&T)[P], float (&cr)[P],
    float (&cg)[P], float (&cb)[P], int32_t (&cur)[P]) {
    constexpr int NT = kTilePix / P;
    const int tid = threadIdx.x, w = tid / kWave;
    // Finished pixels as wave masks (scalar registers): bit l of dm[k] is
    // set once lane l's k-th pixel is done (outside the image, or T <= 1e-4).
    // The per-lane view of a mask is an inverse ballot, which costs no vector
    // instruction, so the bookkeeping stays on the scalar unit.
    uint64_t dm[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        dm[k] = __builtin_amdgcn_ballot_w64(!q.inside[k]);
        T[k] = 1.f;
        cr[k] = cg[k] = cb[k] = 0.f;
        cur[k] = 0;
    }
    auto wave_live = [&]() {
        bool live = false;
#pragma unroll
        for (int k = 0; k < P; ++k) live |= dm[k] != ~0ull;
        return live;
    };
    // The bands' pixel centres (NDC): lane 0 holds each band's first column
    // and top row, lane 63 its last column and bottom row.
    const float wx0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q.px), 0));
    const float wx1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q.px), kWave - 1));
    float wy0[P], wy1[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        wy1[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q.py[k]), 0));
        wy0[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q.py[k]), kWave - 1));
    }
    // the tile's row and column in its image (uniform): with BITS the
    // backward may reuse the last batch, so its pair slots are staged too
    const uint32_t trow = (uint32_t)__builtin_amdgcn_readfirstlane(q.row[0] >> 4);
    const uint32_t tcol = (uint32_t)__builtin_amdgcn_readfirstlane(q.j >> 4);
    int loaded = -1;
    // the list entries (splat ids) of the next batch are read a batch ahead
    uint32_t gnext = range.x + tid < range.y ? a.flat[range.x + tid] : 0u;
    for (int b = range.x; b < range.y; b += NT) {
        if (__syncthreads_count(wave_live()) == 0) break;
        const int idx = b + tid;
        const uint32_t g = gnext;
        if (idx + NT < range.y) gnext = a.flat[idx + NT];
        if (idx < range.y) stage_record<BITS>(a, g, tid, trow, tcol, sA, sB, sC);
        __syncthreads();
        loaded = b;
        const int bs = wave_live() ? min(NT, range.y - b) : 0;
        bool finished = false;
        // 64 splats at a time: one word of reach bits (NT is a multiple of 64)
        for (int t0 = 0; t0 < bs && !finished; t0 += 64) {
            const int t1 = min(bs, t0 + 64);
            // Candidates per band: the splats whose reach ellipse meets the
            // band (no pixel outside it can pass the alpha test); a splat is
            // evaluated only on those bands. Lane l tests splat t0 + l.
            uint64_t cand[P], any = 0;
            {
                const int t = t0 + __lane_id();
                const bool in = t < t1;
                const float4 A = sA[in ? t : t0], B = sB[in ? t : t0];
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    // a band whose pixels are all done has no candidates
                    cand[k] = __ballot(in && reaches_band(A, B, wx0, wx1, wy0[k], wy1[k]));
                    cand[k] = dm[k] == ~0ull ? 0 : cand[k];
                    any |= cand[k];
                }
            }
            uint64_t reach[P];
#pragma unroll
            for (int k = 0; k < P; ++k) reach[k] = 0;
            while (any) {
                const int s = (int)__builtin_ctzll(any);
                any &= any - 1;
                const int t = t0 + s;
                const float4 A = sA[t], B = sB[t], C = sC[t];
                // the colour is read with the record's other parts, through
                // the same LDS address (most candidates reach a pixel)
                asm volatile("" ::"v"(C.x), "v"(C.y), "v"(C.z));
                // Pass 1: which pixels of the candidate bands the splat
                // reaches; a wave whose pixels it reaches none of skips the
                // rest.
                AlphaEval e[P];
                uint64_t okm[P];
                bool anyv = false;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    okm[k] = 0;
                    if ((cand[k] >> s) & 1) {
                        eval_falloff(A, B, q.px, q.py[k], e[k]);
                        okm[k] = alpha_mask(e[k]) & ~dm[k];
                    }
                    anyv |= okm[k] != 0;
                }
                if (!anyv) continue;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    if (okm[k] == 0) continue;  // band k unchanged
                    reach[k] |= 1ull << s;
                    // A skipped pixel gets alpha 0, which leaves T and the
                    // colour exactly unchanged (T * (1 - 0) == T).
                    const bool ok = __builtin_amdgcn_inverse_ballot_w64(okm[k]);
                    const float alpha = ok ? e[k].alpha() : 0.f;
                    const float vis = alpha * T[k];
                    cr[k] += C.x * vis;
                    cg[k] += C.y * vis;
                    cb[k] += C.z * vis;
                    cur[k] = ok ? b + t : cur[k];
                    T[k] = T[k] * (1.f - alpha);
                    dm[k] |= __builtin_amdgcn_ballot_w64(T[k] <= kTThres);
                    if (dm[k] == ~0ull) cand[k] = 0;  // band k done
                }
                if (!wave_live()) {
                    finished = true;
                    break;
                }
                if constexpr (P > 1) {  // drop the candidates of bands done meanwhile
                    uint64_t left = 0;
#pragma unroll
                    for (int k = 0; k < P; ++k) left |= cand[k];
                    any &= left;
                }
            }
            if constexpr (BITS) {
                const int word = (b - range.x + t0) >> 6;
                if (word < kSkipWords && __lane_id() == 0) {
#pragma unroll
                    for (int k = 0; k < P; ++k) s_bits[w][k][word] = reach[k];
                }
            }
        }
    }
    return loaded;
}

// Back to front over the tile's list from each pixel's last contributor
// (compositeBwdKernel), from the final transmittance T, last index lst and
// dL/dC (gr, gg, gb) of each pixel, in the forward's batches of NT splats.
// The batch starting at list index `loaded` is already in sA/sB/sC (-1:
// none). A splat is evaluated band by band (16 x 4 pixels), and its nine
// sums are reduced once over the wave's bands; with BITS, (band, splat)
// pairs whose reach bit the forward cleared are skipped without evaluation.
template <int GRAD_MODE, int P, bool BITS>
__device__ __forceinline__ void composite_back_to_front(
    const CompositeArgs &a, int2 range, const TilePixels<P> &q, float4 *sA, float4 *sB,
    float4 *sC, float (*sG)[kTilePix / P > kWave ? kTilePix / P : 1], int *s_last,
    const uint64_t (*s_bits)[P][kSkipWords], int loaded, float (&T)[P],
    const int32_t (&lst)[P], const float (&gr)[P], const float (&gg)[P], const float (&gb)[P]) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    const int tid = threadIdx.x, w = tid / kWave;
    // the tile's row and column in its image (uniform), and the epoch tag
    const uint32_t trow = (uint32_t)__builtin_amdgcn_readfirstlane(q.row[0] >> 4);
    const uint32_t tcol = (uint32_t)__builtin_amdgcn_readfirstlane(q.j >> 4);
    const uint32_t ep = *a.epoch;
    float acc[P];
    int my_last = -1;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        acc[k] = 0.f;
        my_last = max(my_last, lst[k]);
    }
    // wave- and workgroup-uniform: the splat loop runs on scalar registers
    const int wave_last = __builtin_amdgcn_readfirstlane(wave_max(my_last));
    int block_last = wave_last;
    if constexpr (NW > 1) {
        __syncthreads();  // LDS of a previous phase is free
        if (tid == 0) *s_last = -1;
#pragma unroll
        for (int f = 0; f < 9; ++f) sG[f][tid] = 0.f;
        __syncthreads();
        if (__lane_id() == 0) atomicMax(s_last, wave_last);
        __syncthreads();
        block_last = __builtin_amdgcn_readfirstlane(*s_last);
    }
    // The field this lane holds after the transposed reduction (of its row,
    // or of the wave in row 0), and the factor that turns the sum into the
    // gradient (field_scale).
    const int lane = __lane_id();
    const int my_field = field_of_row_lane(lane);
    const float my_scale = field_scale(my_field);
    const int end = min(range.y, block_last + 1);
    if (end <= range.x) return;  // workgroup-uniform
    for (int bb = range.x + (end - 1 - range.x) / NT * NT; bb >= range.x; bb -= NT) {
        const int bs = min(NT, end - bb);  // slots [0, bs) hold list indices bb + slot
        if (bb != loaded) {
            __syncthreads();
            if (tid < bs) stage_record<true>(a, a.flat[bb + tid], tid, trow, tcol, sA, sB, sC);
            __syncthreads();
        }
        loaded = -1;  // only the forward's last batch is reused
        // 64 slots at a time, last first (bb - range.x is a multiple of 64,
        // so a chunk is one word of reach bits); the set bits of the chunk's
        // mask are visited from the highest down.
        const int tmax = min(bs - 1, wave_last - bb);
        for (int c0 = tmax < 0 ? -64 : tmax & ~63; c0 >= 0; c0 -= 64) {
            const int lim = tmax - c0;  // highest slot of the chunk in range
            const uint64_t mask = lim >= 63 ? ~0ull : (2ull << lim) - 1ull;
            // per band: the chunk's slots to visit (the forward's reach bits
            // where recorded), and their union
            uint64_t mk[P], any = 0;
#pragma unroll
            for (int k = 0; k < P; ++k) {
                mk[k] = mask;
                if constexpr (BITS) {
                    const int word = (bb - range.x + c0) >> 6;
                    if (word < kSkipWords) {
                        const uint64_t v = s_bits[w][k][word];
                        mk[k] &= ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32))
                                  << 32) |
                                 __builtin_amdgcn_readfirstlane((uint32_t)v);
                    }
                }
                any |= mk[k];
            }
            while (any) {
                const int sl = 63 - (int)__builtin_clzll(any);
                any &= ~(1ull << sl);
                const int t = c0 + sl;
                const int idx = bb + t;
                const float4 A = sA[t], B = sB[t], C = sC[t];
                asm volatile("" ::"v"(C.x), "v"(C.y), "v"(C.z));  // read with A and B
                // sums start at -0, the identity of float addition (+0 is not:
                // +0 + -0 = +0), so the first band's terms need no add
                float v[9];
#pragma unroll
                for (int f = 0; f < 9; ++f) v[f] = -0.f;
                bool anyk = false;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    if (!((mk[k] >> sl) & 1)) continue;  // band k not reached
                    // which of the band's pixels the splat reaches; a band it
                    // reaches none of is skipped
                    AlphaEval e;
                    const bool ok = eval_alpha(A, B, q.px, q.py[k], e) & (idx <= lst[k]);
                    if (__ballot(ok) == 0) continue;
                    anyk = true;
                    // without branches: a skipped pixel gets alpha 0 and gate
                    // false, which adds exact zeros and keeps T and acc unchanged
                    const float alpha = ok ? e.alpha() : 0.f;
                    // 1 / max(1 - alpha, 1e-6) of the reference: alpha <= 0.99, so the
                    // floor never applies
                    const float ra = __builtin_amdgcn_rcpf(1.f - alpha);
                    // transmittance in front of the splat (a skipped pixel: alpha 0,
                    // ra = 1 exactly, T unchanged)
                    const float Tb = T[k] * ra;
                    const float fac = alpha * Tb;
                    v[5] = fmaf(fac, gr[k], v[5]);
                    v[6] = fmaf(fac, gg[k], v[6]);
                    v[7] = fmaf(fac, gb[k], v[7]);
                    const float cgk = fmaf(C.x, gr[k], fmaf(C.y, gg[k], C.z * gb[k]));
                    float v_alpha;
                    if constexpr (GRAD_MODE == DS_GRAD_REFERENCE)
                        v_alpha = Tb * (cgk - acc[k]);
                    else
                        v_alpha = Tb * cgk - ra * acc[k];
                    const bool gate = ok & (e.og <= kMaxAlpha);  // alpha below its clamp
                    const float gva = (gate ? e.g : 0.f) * v_alpha;
                    v[8] += gva;
                    // m = -2 dL/d dist2 (dL/d dist2 = -o g v_alpha / 2); the
                    // per-field factors are applied once, after the reduction
                    const float m = B.y * gva;
                    const float pu = m * e.u, pw = m * e.w;
                    v[0] += pu;
                    v[1] += pw;
                    v[2] = fmaf(pu, e.u, v[2]);
                    v[3] = fmaf(pu, e.w, v[3]);
                    v[4] = fmaf(pw, e.w, v[4]);
                    acc[k] = fmaf(fac, cgk, acc[k]);
                    T[k] = Tb;
                }
                if (!anyk) continue;
                if constexpr (NW == 1) {
                    // the wave's sums are the tile's: nine lanes of row 0
                    // store the fields, lane 6 the epoch tag
                    const float sum = my_scale * wave_sum9_transposed(v, lane);
                    const uint32_t slot =
                        (uint32_t)__builtin_amdgcn_readfirstlane(__float_as_uint(C.w));
                    float *r = reinterpret_cast<float *>(a.pair_rec + 3 * (size_t)slot);
                    if (lane < 16 && my_field >= 0) r[my_field] = sum;
                    if (lane == 6) r[9] = __uint_as_float(ep);
                } else {
                    // the tile's waves add their sums in LDS (nine lanes of
                    // row 0, one field each); the field factors are applied
                    // when the slot's record is written. (Adding the four
                    // rows' sums there instead of shuffling them together
                    // was 40-50% slower: same-address LDS atomics serialise.)
                    const float sum = wave_sum9_transposed(v, lane);
                    if (lane < 16 && my_field >= 0) atomicAdd(&sG[my_field][t], sum);
                }
            }
        }
        if constexpr (NW > 1) {
            __syncthreads();
            if (tid < bs) {  // the slot's record: one thread's
                float gs[9];
                bool hit = false;
#pragma unroll
                for (int f = 0; f < 9; ++f) {
                    gs[f] = sG[f][tid];
                    hit |= gs[f] != 0.f;
                }
                // A slot no pixel reached keeps its sums at zero and its
                // record unwritten (an older epoch, which counts as zero).
                if (hit) {
                    float4 *r = a.pair_rec + 3 * (size_t)__float_as_uint(sC[tid].w);
                    r[0] = make_float4(gs[0] * field_scale(0), gs[1] * field_scale(1),
                                       gs[2] * field_scale(2), gs[3] * field_scale(3));
                    r[1] = make_float4(gs[4] * field_scale(4), gs[5], gs[6], gs[7]);
                    r[2] = make_float4(gs[8], __uint_as_float(ep), 0.f, 0.f);
#pragma unroll
                    for (int f = 0; f < 9; ++f) sG[f][tid] = 0.f;
                }
            }
        }
    }
}

// The slots' records of this epoch summed per splat and added to its
// gradient. A splat's slots are contiguous (slot_splat holds each slot's
// splat): a run. Each workgroup takes spans of kReduceSpan consecutive slots,
// one 64-slot window per wave, one slot per lane, and adds them with a
// segmented inclusive scan over the lanes: shifts up by 1, 2, 4 and 8 inside
// each 16-lane row (DPP), each adding only within the run, leave every lane
// the sum of its run from the row's start; then the odd rows add the
// previous row's last lane and rows 2 and 3 add lane 31, each when the run
// continues across that row boundary. A run going on from an earlier window
// of the span gets those windows' sums through LDS (each wave posts its last
// lane's sum and splat). The last lane of each run's part in the span writes
// the sums: a plain store (or add) for a run that lies inside the span (no
// other workgroup touches it), a float atomic for a run that crosses the
// span's start or end (at most two per span; device-scope float atomics on
// scattered words are slow, which is why a span is longer than one window).
// With STORE the gradient is known to be zero (cleared by the same step), so a
// whole run is stored without reading it.
constexpr int kSpanWindows = 4;  // 64-slot windows per span, one per wave
constexpr uint32_t kReduceSpan = kSpanWindows * kWave;
constexpr uint32_t kNoSplat = 0xFFFFFFFFu;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}

// One in-row step: lane l adds lane l - D of its row when both hold splat g.
template <int D>
__device__ __forceinline__ void seg_scan_row_step(float (&v)[9], uint32_t g, int lane) {
    constexpr int kRowShr = 0x110 + D;  // row_shr:D (lanes below D of a row read 0)
    const uint32_t gu = dpp_u32<kRowShr>(g);
    const bool same = (lane & 15) >= D && gu == g;
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        const float u = __uint_as_float(dpp_u32<kRowShr>(__float_as_uint(v[f])));
        v[f] += same ? u : 0.f;
    }
}

// Lane l adds lane src's sums when take is set (src read by every lane).
__device__ __forceinline__ void seg_scan_carry(float (&v)[9], int src, bool take) {
#pragma unroll
    for (int f = 0; f < 9; ++f) {
        const float u = __shfl(v[f], src, kWave);
        v[f] += take ? u : 0.f;
    }
}

// 64 consecutive slots of a wave: each lane's slot splat and record, and the
// splat of the slot after the last.
struct SlotLoad {
    uint32_t g, g_next;
    float4 r0, r1, r2;
};

__device__ __forceinline__ SlotLoad load_slots(const float4 *__restrict__ rec,
                                               const uint32_t *__restrict__ slot_splat,
                                               uint32_t base, uint32_t m, int lane) {
    SlotLoad L;
    const uint32_t p = base + lane;
    const bool valid = p < m;
    L.g = valid ? slot_splat[p] : kNoSplat;
    L.g_next = base + kWave < m ? slot_splat[base + kWave] : kNoSplat;
    const size_t q = valid ? 3 * (size_t)p : 0;
    L.r0 = rec[q];
    L.r1 = rec[q + 1];
    L.r2 = rec[q + 2];
    return L;
}

// The next forward's dispatch order (NextOrder) by one workgroup: pass 1
// buckets every tile (kept in LDS) and counts the buckets per wave (the lanes
// of one bucket found with five ballots, their first lane adding their
// number to the wave's row); a descending scan gives each bucket's first
// position; pass 2 reserves each wave's positions with one LDS atomic per
// bucket.
constexpr int kOrderLoads = 8;  // ranges in flight per thread
__device__ __forceinline__ void next_order_block(const NextOrder o) {
    constexpr int NT = kReduceSpan, NW = NT / kWave;
    __shared__ uint32_t h[32];
    __shared__ uint32_t wh[NW][32];
    __shared__ uint8_t s_b[kNextOrderMaxTiles];
    const int tid = threadIdx.x, w = tid / kWave;
    const uint64_t lt = lanemask_lt();
    for (int i = tid; i < NW * 32; i += NT) (&wh[0][0])[i] = 0u;
    __syncthreads();
    for (int t0 = 0; t0 < o.ntiles; t0 += NT * kOrderLoads) {
        int2 r[kOrderLoads];
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const int t = t0 + u * NT + tid;
            r[u] = t < o.ntiles ? o.ranges[t] : make_int2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const int t = t0 + u * NT + tid;
            const bool valid = t < o.ntiles;
            const uint32_t b = tile_bucket(r[u]);
            if (valid) s_b[t] = (uint8_t)b;
            const uint64_t peers = same_bucket(b, valid);
            if (valid && (peers & lt) == 0) wh[w][b] += (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
    if (tid < 32) {  // bucket totals
        uint32_t v = 0;
        for (int q = 0; q < NW; ++q) v += wh[q][tid];
        h[tid] = v;
    }
    __syncthreads();
    if (tid == 0) {  // descending exclusive scan: the first position of each bucket
        uint32_t run = 0;
        for (int b = 31; b >= 0; --b) {
            const uint32_t v = h[b];
            h[b] = run;
            run += v;
        }
    }
    __syncthreads();
    for (int t0 = 0; t0 < o.ntiles; t0 += NT) {
        const int t = t0 + tid;
        const bool valid = t < o.ntiles;
        const uint32_t b = valid ? (uint32_t)s_b[t] : 0u;
        const uint64_t peers = same_bucket(b, valid);
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        uint32_t base = 0;
        if (valid && rank == 0) base = atomicAdd(&h[b], (uint32_t)__popcll(peers));
        const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : 0;
        base = (uint32_t)__shfl((int)base, leader, kWave);
        if (valid) o.inv[t] = base + rank;
    }
}

template <bool STORE>
__global__ __launch_bounds__(kReduceSpan) void k_pair_reduce(
    const float4 *__restrict__ rec, const uint32_t *__restrict__ slot_splat,
    const uint32_t *nisect, const uint32_t *epoch, ds_splat2d grad, NextOrder next) {
    // with a next order to compute, workgroup 0 does that and the spans
    // start at workgroup 1
    const uint32_t first = next.ranges ? 1u : 0u;
    if (blockIdx.x < first) {
        next_order_block(next);
        return;
    }
    // per wave of the span: its last lane's splat, whether that lane's run
    // covers the whole window, and its sums
    __shared__ uint32_t s_g[kSpanWindows];
    __shared__ uint32_t s_whole[kSpanWindows];
    __shared__ float s_sum[kSpanWindows][9];
    const uint32_t m = *nisect, ep = *epoch;
    const int lane = __lane_id(), w = threadIdx.x / kWave;
    float *const dst[9] = {grad.pos_x,  grad.pos_y,   grad.cov_xx,  grad.cov_xy, grad.cov_yy,
                           grad.color_r, grad.color_g, grad.color_b, grad.opacity};
    for (uint32_t s0 = (blockIdx.x - first) * kReduceSpan; s0 < m;
         s0 += (gridDim.x - first) * kReduceSpan) {  // workgroup-uniform
        const uint32_t s1 = min(s0 + kReduceSpan, m);
        const uint32_t base = s0 + w * kWave;
        // the runs crossing the span's start and its end (if any)
        const uint32_t g_lead = s0 > 0 ? slot_splat[s0 - 1] : kNoSplat;
        const uint32_t g_after = s1 < m ? slot_splat[s1] : kNoSplat;
        const bool live = base < s1;  // wave-uniform
        uint32_t g = kNoSplat;
        float v[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) v[f] = 0.f;
        bool valid = false, last = false, atomic = false, plain = false;
        float old[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) old[f] = 0.f;
        if (live) {
            const SlotLoad cur = load_slots(rec, slot_splat, base, m, lane);
            const uint32_t p = base + lane;
            valid = p < m;
            g = cur.g;
            const bool on = valid && __float_as_uint(cur.r2.y) == ep;
            v[0] = on ? cur.r0.x : 0.f;
            v[1] = on ? cur.r0.y : 0.f;
            v[2] = on ? cur.r0.z : 0.f;
            v[3] = on ? cur.r0.w : 0.f;
            v[4] = on ? cur.r1.x : 0.f;
            v[5] = on ? cur.r1.y : 0.f;
            v[6] = on ? cur.r1.z : 0.f;
            v[7] = on ? cur.r1.w : 0.f;
            v[8] = on ? cur.r2.x : 0.f;
            // the last lane of a run's part in the span (the next slot holds
            // another splat or lies past the span) writes it: atomically if
            // the run crosses the span's start or end
            const uint32_t gn = __shfl_down(g, 1, kWave);
            last = p + 1 == s1 || (lane == kWave - 1 ? cur.g_next != g : gn != g);
            atomic = g == g_lead || (p + 1 == s1 && g == g_after);
            plain = valid && last && !atomic;
            const uint32_t gp = __shfl_up(g, 1, kWave);
            const uint64_t firsts = __ballot(lane == 0 || gp != g);
            if constexpr (!STORE) {
                // adding to the gradient: its old values load during the scan,
                // for the runs with a record of this epoch or an earlier
                // window's part
                const uint64_t onm = __ballot(on);
                const uint64_t upto = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1ull;
                const int st = 63 - (int)__builtin_clzll(firsts & upto);
                const bool any = ((onm & upto) >> st) != 0 || st == 0;
                if (plain && any) {
#pragma unroll
                    for (int f = 0; f < 9; ++f) old[f] = dst[f][g];
                }
            }
            // runs of equal g are contiguous: lanes a < b hold the same splat
            // exactly when every lane between does
            seg_scan_row_step<1>(v, g, lane);
            seg_scan_row_step<2>(v, g, lane);
            seg_scan_row_step<4>(v, g, lane);
            seg_scan_row_step<8>(v, g, lane);
            {  // rows 1 and 3: the run's part in rows 0 and 2 (lanes 15, 47)
                const int row0 = lane & ~15;
                const int src = row0 > 0 ? row0 - 1 : 0;
                const uint32_t g_src = __shfl(g, src, kWave), g_row0 = __shfl(g, row0, kWave);
                seg_scan_carry(v, src, ((lane >> 4) & 1) && g_src == g && g_row0 == g);
            }
            {  // rows 2 and 3: the run's part in rows 0 and 1 (lane 31)
                const uint32_t g31 = __shfl(g, 31, kWave), g32 = __shfl(g, 32, kWave);
                seg_scan_carry(v, 31, lane >= 32 && g31 == g && g32 == g);
            }
            if (lane == kWave - 1) {
                s_g[w] = g;
                s_whole[w] = firsts == 1ull;  // one run over the whole window
#pragma unroll
                for (int f = 0; f < 9; ++f) s_sum[w][f] = v[f];
            }
        }
        __syncthreads();
        if (live) {
            // the part of this window's first run in the span's earlier windows
            const uint32_t g0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
            float carry[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) carry[f] = 0.f;
            for (int k = w - 1; k >= 0 && s_g[k] == g0; --k) {  // uniform
#pragma unroll
                for (int f = 0; f < 9; ++f) carry[f] += s_sum[k][f];
                if (!s_whole[k]) break;
            }
#pragma unroll
            for (int f = 0; f < 9; ++f) v[f] += g == g0 ? carry[f] : 0.f;
            if (plain) {
#pragma unroll
                for (int f = 0; f < 9; ++f)
                    if (v[f] != 0.f) dst[f][g] = STORE ? v[f] : old[f] + v[f];
            } else if (valid && last) {
#pragma unroll
                for (int f = 0; f < 9; ++f)
                    if (v[f] != 0.f) atomicAdd(dst[f] + g, v[f]);
            }
        }
        __syncthreads();  // the LDS is reused by the next span
    }
}

// The MSE target of this thread's pixels, loaded when the kernel starts so
// that its latency passes while the tile composites (zero outside the image
// or without a target).
template <bool MSE, int P>
struct PixelTarget {
    float r[P], g[P], b[P];
};

template <bool MSE, int P>
__device__ __forceinline__ PixelTarget<MSE, P> load_target(const TileImage &im,
                                                           const TilePixels<P> &q,
                                                           const float *__restrict__ target) {
    PixelTarget<MSE, P> t;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        t.r[k] = t.g[k] = t.b[k] = 0.f;
        if constexpr (MSE) {
            if (q.inside[k]) {
                const int pix = im.pix_base + q.row[k] * im.W + q.j;
                t.r[k] = target[3 * pix + 0];
                t.g[k] = target[3 * pix + 1];
                t.b[k] = target[3 * pix + 2];
            }
        }
    }
    return t;
}

template <bool MSE, int P>
__device__ __forceinline__ void composite_epilogue(
    const TileImage &im, const TilePixels<P> &q, const float (&T)[P], const float (&cr)[P],
    const float (&cg)[P], const float (&cb)[P], const int32_t (&cur)[P],
    float *__restrict__ out_rgb, float *__restrict__ out_T, int32_t *__restrict__ out_last,
    const PixelTarget<MSE, P> &tg, float *loss, float *__restrict__ grad, float *s_red,
    float (&gr)[P], float (&gg)[P], float (&gb)[P]) {
    constexpr int NW = kTilePix / P / kWave;
    const int tid = threadIdx.x;
    float l = 0.f;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        gr[k] = gg[k] = gb[k] = 0.f;
        if (!q.inside[k]) continue;
        const int pix = im.pix_base + q.row[k] * im.W + q.j;
        out_rgb[3 * pix + 0] = cr[k];
        out_rgb[3 * pix + 1] = cg[k];
        out_rgb[3 * pix + 2] = cb[k];
        out_T[pix] = T[k];
        out_last[pix] = cur[k];
        if constexpr (MSE) {
            const float dr = cr[k] - tg.r[k];
            const float dg = cg[k] - tg.g[k];
            const float db = cb[k] - tg.b[k];
            const float s2 = 2.f * im.inv_npix;
            gr[k] = s2 * dr;
            gg[k] = s2 * dg;
            gb[k] = s2 * db;
            grad[3 * pix + 0] = gr[k];
            grad[3 * pix + 1] = gg[k];
            grad[3 * pix + 2] = gb[k];
            l += (dr * dr + dg * dg + db * db) * im.inv_npix;
        }
    }
    if constexpr (MSE) {
        l = wave_sum(l);
        if (__lane_id() == 0) s_red[tid / kWave] = l;
        __syncthreads();
        if (tid == 0) {
            float s = 0.f;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += s_red[w];
            atomicAdd(loss + im.inst, s);
        }
    }
}

template <bool MSE, int P, bool MULTI>
__global__ __launch_bounds__(kTilePix / P) void k_composite_fwd(
    CompositeArgs a, float *__restrict__ out_rgb, float *__restrict__ out_T,
    int32_t *__restrict__ out_last, const float *__restrict__ target, float *loss,
    float *__restrict__ grad, float inv_npix_one) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    __shared__ float4 sA[NT], sB[NT], sC[NT];
    __shared__ float s_red[NW];
    const int tile = tile_of_block(a);
    const TileImage im = tile_image<MULTI>(a, tile, inv_npix_one);
    const TilePixels<P> q = tile_pixels<P>(im, threadIdx.x);
    const int2 range = range_of_block(a);
    const PixelTarget<MSE, P> tg = load_target<MSE, P>(im, q, target);
    float T[P], cr[P], cg[P], cb[P], gr[P], gg[P], gb[P];
    int32_t cur[P];
    (void)composite_front_to_back<P, false>(a, range, q, sA, sB, sC, nullptr, T, cr, cg, cb,
                                            cur);
    composite_epilogue<MSE, P>(im, q, T, cr, cg, cb, cur, out_rgb, out_T, out_last, tg, loss,
                               grad, s_red, gr, gg, gb);
}

template <int GRAD_MODE, int P, bool MULTI>
__global__ __launch_bounds__(kTilePix / P) void k_composite_bwd(
    CompositeArgs a, const float *__restrict__ Tfin, const int32_t *__restrict__ last,
    const float *__restrict__ gout) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    constexpr int GW = NW > 1 ? NT : 1;
    __shared__ float4 sA[NT], sB[NT], sC[NT];
    __shared__ float sG[9][GW];
    __shared__ int s_last;
    const int tile = tile_of_block(a);
    const int2 range = range_of_block(a);
    if (range.y <= range.x) return;
    const TileImage im = tile_image<MULTI>(a, tile, 0.f);
    const TilePixels<P> q = tile_pixels<P>(im, threadIdx.x);
    // Per pixel: transmittance (walked back from T_final), last contributor,
    // dL/dC.
    float T[P], gr[P], gg[P], gb[P];
    int32_t lst[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int pix = q.inside[k] ? im.pix_base + q.row[k] * im.W + q.j : 0;
        T[k] = q.inside[k] ? Tfin[pix] : 1.f;
        lst[k] = q.inside[k] ? last[pix] : -1;
        gr[k] = q.inside[k] ? gout[3 * pix + 0] : 0.f;
        gg[k] = q.inside[k] ? gout[3 * pix + 1] : 0.f;
        gb[k] = q.inside[k] ? gout[3 * pix + 2] : 0.f;
    }
    composite_back_to_front<GRAD_MODE, P, false>(a, range, q, sA, sB, sC, sG, &s_last,
                                                 nullptr, -1, T, lst, gr, gg, gb);
}

// Forward with the fused MSE, then the backward of the same tile in the same
// workgroup: the per-pixel state (T, last index, dL/dC) stays in registers
// and the backward skips the (band, splat) pairs the forward found unreached.
template <int GRAD_MODE, int P, bool MULTI>
__global__ __launch_bounds__(kTilePix / P) void k_composite_fused(
    CompositeArgs a, float *__restrict__ out_rgb, float *__restrict__ out_T,
    int32_t *__restrict__ out_last, const float *__restrict__ target, float *loss,
    float *__restrict__ grad_rgb, float inv_npix_one) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    constexpr int GW = NW > 1 ? NT : 1;
    __shared__ float4 sA[NT], sB[NT], sC[NT];
    __shared__ float sG[9][GW];
    __shared__ float s_red[NW];
    __shared__ int s_last;
    __shared__ uint64_t s_bits[NW][P][kSkipWords];
    const int tile = tile_of_block(a);
    const TileImage im = tile_image<MULTI>(a, tile, inv_npix_one);
    const TilePixels<P> q = tile_pixels<P>(im, threadIdx.x);
    const int2 range = range_of_block(a);
    const PixelTarget<true, P> tg = load_target<true, P>(im, q, target);
    float T[P], cr[P], cg[P], cb[P], gr[P], gg[P], gb[P];
    int32_t cur[P];
    const int loaded =
        composite_front_to_back<P, true>(a, range, q, sA, sB, sC, s_bits, T, cr, cg, cb, cur);
    composite_epilogue<true, P>(im, q, T, cr, cg, cb, cur, out_rgb, out_T, out_last, tg,
                                loss, grad_rgb, s_red, gr, gg, gb);
    if (range.y <= range.x) return;
    int32_t lst[P];
#pragma unroll
    for (int k = 0; k < P; ++k) lst[k] = q.inside[k] ? cur[k] : -1;
    composite_back_to_front<GRAD_MODE, P, true>(a, range, q, sA, sB, sC, sG, &s_last,
                                                s_bits, loaded, T, lst, gr, gg, gb);
}

}  // namespace

#define DS_PIX_DISPATCH(P, ...)                               \
    switch (P) {                                              \
        case 1: { constexpr int PP = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int PP = 2; __VA_ARGS__; } break; \
        default: { constexpr int PP = 4; __VA_ARGS__; } break; \
    }

template <bool MULTI>
static void composite_fwd(const CompositeArgs &a, int P, float *out_rgb, float *out_T,
                          int32_t *out_last, const float *target, float *loss, float *grad,
                          hipStream_t s) {
    const float inv_npix = 1.f / ((float)a.W * (float)a.H);
    if (target) {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_fwd<true, PP, MULTI>),
                                              dim3(a.ntiles), dim3(kTilePix / PP), 0, s, a,
                                              out_rgb, out_T, out_last, target, loss, grad,
                                              inv_npix));
    } else {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_fwd<false, PP, MULTI>),
                                              dim3(a.ntiles), dim3(kTilePix / PP), 0, s, a,
                                              out_rgb, out_T, out_last, (const float *)nullptr,
                                              (float *)nullptr, (float *)nullptr, inv_npix));
    }
}

template <int MODE, bool MULTI>
static void composite_bwd(const CompositeArgs &a, int P, const float *T, const int32_t *last,
                          const float *grad_rgb, hipStream_t s) {
    DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_bwd<MODE, PP, MULTI>), dim3(a.ntiles),
                                          dim3(kTilePix / PP), 0, s, a, T, last, grad_rgb));
}

template <int MODE, bool MULTI>
static void composite_fused(const CompositeArgs &a, int P, float *out_rgb, float *out_T,
                            int32_t *out_last, const float *target, float *loss, float *grad_rgb,
                            hipStream_t s) {
    const float inv_npix = 1.f / ((float)a.W * (float)a.H);
    DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_fused<MODE, PP, MULTI>), dim3(a.ntiles),
                                          dim3(kTilePix / PP), 0, s, a, out_rgb, out_T, out_last,
                                          target, loss, grad_rgb, inv_npix));
}

ds_status launch_composite_fused(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                                 float *out_rgb, float *out_T, int32_t *out_last,
                                 const float *target, float *loss, float *grad_rgb,
                                 hipStream_t s) {
    if (a.ntiles <= 0) return DS_OK;
    const int P = pix_per_thread;
    if (grad_mode == DS_GRAD_REFERENCE) {
        if (a.mi)
            composite_fused<DS_GRAD_REFERENCE, true>(a, P, out_rgb, out_T, out_last, target, loss,
                                                     grad_rgb, s);
        else
            composite_fused<DS_GRAD_REFERENCE, false>(a, P, out_rgb, out_T, out_last, target,
                                                      loss, grad_rgb, s);
    } else {
        if (a.mi)
            composite_fused<DS_GRAD_EXACT, true>(a, P, out_rgb, out_T, out_last, target, loss,
                                                 grad_rgb, s);
        else
            composite_fused<DS_GRAD_EXACT, false>(a, P, out_rgb, out_T, out_last, target, loss,
                                                  grad_rgb, s);
    }
    DS_CHECK_LAUNCH("composite_fused");
    return DS_OK;
}

ds_status launch_composite_fwd(const CompositeArgs &a, int pix_per_thread, float *out_rgb,
                               float *out_T, int32_t *out_last, const float *target,
                               float *loss, float *grad, hipStream_t s) {
    if (a.ntiles <= 0) return DS_OK;
    if (a.mi)
        composite_fwd<true>(a, pix_per_thread, out_rgb, out_T, out_last, target, loss, grad, s);
    else
        composite_fwd<false>(a, pix_per_thread, out_rgb, out_T, out_last, target, loss, grad, s);
    DS_CHECK_LAUNCH("composite_fwd");
    return DS_OK;
}

ds_status launch_pair_reduce(const float4 *pair_rec, const uint32_t *slot_splat,
                             const uint32_t *nisect, const uint32_t *epoch, int64_t slot_cap,
                             const ds_splat2d &grad, bool grad_is_zero, hipStream_t s,
                             NextOrder next) {
    if (next.ranges && next.ntiles > kNextOrderMaxTiles)
        return set_error(DS_ERR_INVALID_ARG, "next dispatch order over more than 4096 tiles");
    if (slot_cap <= 0 && !next.ranges) return DS_OK;
    // a workgroup per span of kReduceSpan slots up to 16384 workgroups, each
    // then taking the span a grid further on
    const int64_t spans = (std::max<int64_t>(slot_cap, 1) + kReduceSpan - 1) / kReduceSpan;
    const int blocks = (int)std::min<int64_t>(spans, 16384) + (next.ranges ? 1 : 0);
    if (grad_is_zero)
        hipLaunchKernelGGL(k_pair_reduce<true>, dim3(blocks), dim3(kReduceSpan), 0, s, pair_rec,
                           slot_splat, nisect, epoch, grad, next);
    else
        hipLaunchKernelGGL(k_pair_reduce<false>, dim3(blocks), dim3(kReduceSpan), 0, s,
                           pair_rec, slot_splat, nisect, epoch, grad, next);
    DS_CHECK_LAUNCH("pair_reduce");
    return DS_OK;
}

ds_status launch_composite_bwd(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                               const float *T, const int32_t *last, const float *grad_rgb,
                               hipStream_t s) {
    if (a.ntiles <= 0) return DS_OK;
    const int P = pix_per_thread;
    if (grad_mode == DS_GRAD_REFERENCE) {
        if (a.mi)
            composite_bwd<DS_GRAD_REFERENCE, true>(a, P, T, last, grad_rgb, s);
        else
            composite_bwd<DS_GRAD_REFERENCE, false>(a, P, T, last, grad_rgb, s);
    } else {
        if (a.mi)
            composite_bwd<DS_GRAD_EXACT, true>(a, P, T, last, grad_rgb, s);
        else
            composite_bwd<DS_GRAD_EXACT, false>(a, P, T, last, grad_rgb, s);
    }
    DS_CHECK_LAUNCH("composite_bwd");
    return DS_OK;
}

}  // namespace ds
