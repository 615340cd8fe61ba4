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
// MI355X layout: a workgroup owns one tile and each thread P pixels of it
// (P = 4: one wave64 per tile; P = 2: two waves; P = 1: four waves). Splat
// records are staged in LDS in batches of one record per thread and read as
// broadcasts. In the backward the nine per-splat gradient sums are reduced
// across the wave with DPP row operations (no LDS traffic), amortised over P
// pixels per lane, and lane 63 issues one global float atomic per field
// (through an LDS pre-reduction when the tile has more than one wave).
// The transmittance is stored directly (not as 1 - T) so the backward's
// division chain starts from an exact value.
#include "composite.hpp"

namespace ds {
namespace {

// Workgroup -> tile so that the tiles of one XCD (dispatch is round-robin
// over 8 XCDs) form one contiguous row-major band sharing that XCD's L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
    const int q = nblk / 8, r = nblk % 8;
    const int x = bid % 8, local = bid / 8;
    return x * q + min(x, r) + local;
}

// u = (Sigma^-1 d)_x, w = (Sigma^-1 d)_y for d = splat centre - pixel, so that
// dist2 = dx u + dy w, d dist2/d(x,y) = 2 (u, w), and the conic gradients of
// the reference (-1/4 ddx^2, -1/4 ddy^2, -2 b dist2 - 2 dx dy / det) are
// -u^2, -w^2 and -2 u w (the last by u w = b dist2 + dx dy / det).
struct AlphaEval {
    float u, w, dist2, g, og, alpha;
};

// Gaussian falloff and alpha of one splat at one pixel. Forward and backward
// run this same code so their discrete decisions agree bit for bit.
// Returns false when the splat is skipped (sigma < 0 or alpha < 1/255).
__device__ __forceinline__ bool eval_alpha(const float4 &A, const float4 &B, float px, float py,
                                           AlphaEval &e) {
    constexpr float kNegHalfLog2e = -0.72134752044448170f;  // -0.5 / ln 2
    const float dx = A.x - px;
    const float dy = A.y - py;
    e.u = fmaf(A.z, dx, A.w * dy);
    e.w = fmaf(A.w, dx, B.x * dy);
    e.dist2 = fmaf(dx, e.u, dy * e.w);
    if (e.dist2 < 0.f) return false;  // sigma = dist2 / 2 < 0
    e.g = __builtin_amdgcn_exp2f(e.dist2 * kNegHalfLog2e);
    e.og = B.y * e.g;
    e.alpha = fminf(kMaxAlpha, e.og);
    return e.alpha >= kAlphaThres;
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

// Sum over the 64 lanes of a wave, valid in lane 63 only (other lanes hold
// partial sums). All lanes must be active.
__device__ __forceinline__ float wave_sum_to_63(float v) {
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]  -> every lane: its quad's sum
    v += dpp<0x114>(v);  // row_shr:4
    v += dpp<0x118>(v);  // row_shr:8            -> lane 15 of each row: the row's sum
    v += dpp<0x142>(v);  // row_bcast:15         -> lane 63: rows 2 + 3
    v += dpp<0x143>(v);  // row_bcast:31         -> lane 63: all four rows
    return v;
}

template <bool MSE, int P>
__global__ __launch_bounds__(kTilePix / P) void k_composite_fwd(
    CompositeArgs a, float *__restrict__ out_rgb, float *__restrict__ out_T,
    int32_t *__restrict__ out_last, const float *__restrict__ target, float *loss,
    float *__restrict__ grad, float inv_npix) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    constexpr int RSTEP = kTile / P;  // rows between a thread's pixels
    __shared__ float4 sA[NT], sB[NT], sC[NT];
    __shared__ float s_red[NW];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int trow = tile / a.ntx, tcol = tile % a.ntx;
    const int j = tcol * kTile + (tid & 15);
    const float px = pixel_ndc_x(j, a.W);
    float py[P], T[P], cr[P], cg[P], cb[P];
    int32_t cur[P];
    bool done[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int i = trow * kTile + (tid >> 4) + k * RSTEP;
        py[k] = pixel_ndc_y(i, a.H);
        done[k] = !(i < a.H && j < a.W);
        T[k] = 1.f;
        cr[k] = cg[k] = cb[k] = 0.f;
        cur[k] = 0;
    }
    const int2 range = a.ranges[tile];
    for (int b = range.x; b < range.y; b += NT) {
        bool live = false;
#pragma unroll
        for (int k = 0; k < P; ++k) live |= !done[k];
        if (__syncthreads_count(live) == 0) break;
        const int idx = b + tid;
        if (idx < range.y) {
            const float4 *r = a.rec + 3 * (size_t)a.flat[idx];
            sA[tid] = r[0];
            sB[tid] = r[1];
            sC[tid] = r[2];
        }
        __syncthreads();
        const int bs = min(NT, range.y - b);
        for (int t = 0; t < bs; ++t) {
            bool any = false;
#pragma unroll
            for (int k = 0; k < P; ++k) any |= !done[k];
            if (!any) break;
            const float4 A = sA[t], B = sB[t];
#pragma unroll
            for (int k = 0; k < P; ++k) {
                if (done[k]) continue;
                AlphaEval e;
                if (!eval_alpha(A, B, px, py[k], e)) continue;
                const float vis = e.alpha * T[k];
                const float4 C = sC[t];
                cr[k] += C.x * vis;
                cg[k] += C.y * vis;
                cb[k] += C.z * vis;
                cur[k] = b + t;
                T[k] = T[k] * (1.f - e.alpha);
                if (T[k] <= kTThres) done[k] = true;
            }
        }
    }
    float l = 0.f;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int i = trow * kTile + (tid >> 4) + k * RSTEP;
        if (!(i < a.H && j < a.W)) continue;
        const int pix = i * a.W + j;
        out_rgb[3 * pix + 0] = cr[k];
        out_rgb[3 * pix + 1] = cg[k];
        out_rgb[3 * pix + 2] = cb[k];
        out_T[pix] = T[k];
        out_last[pix] = cur[k];
        if constexpr (MSE) {
            const float dr = cr[k] - target[3 * pix + 0];
            const float dg = cg[k] - target[3 * pix + 1];
            const float db = cb[k] - target[3 * pix + 2];
            const float s2 = 2.f * inv_npix;
            grad[3 * pix + 0] = s2 * dr;
            grad[3 * pix + 1] = s2 * dg;
            grad[3 * pix + 2] = s2 * db;
            l += (dr * dr + dg * dg + db * db) * inv_npix;
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
            atomicAdd(loss, s);
        }
    }
}

template <int GRAD_MODE, int P>
__global__ __launch_bounds__(kTilePix / P) void k_composite_bwd(
    CompositeArgs a, const float *__restrict__ Tfin, const int32_t *__restrict__ last,
    const float *__restrict__ gout, ds_splat2d grad) {
    constexpr int NT = kTilePix / P;
    constexpr int NW = NT / kWave;
    constexpr int RSTEP = kTile / P;
    constexpr int GW = NW > 1 ? NT : 1;
    __shared__ float4 sA[NT], sB[NT], sC[NT];
    __shared__ float sG[9][GW];
    __shared__ int s_last;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int2 range = a.ranges[tile];
    if (range.y <= range.x) return;
    const int tid = threadIdx.x;
    const int trow = tile / a.ntx, tcol = tile % a.ntx;
    const int j = tcol * kTile + (tid & 15);
    const float px = pixel_ndc_x(j, a.W);
    // Per pixel: transmittance (walked back from T_final), last contributor,
    // dL/dC, and acc = (colour of the splats behind) . dL/dC.
    float py[P], T[P], gr[P], gg[P], gb[P], acc[P];
    int32_t lst[P];
    int my_last = -1;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int i = trow * kTile + (tid >> 4) + k * RSTEP;
        py[k] = pixel_ndc_y(i, a.H);
        const bool inside = (i < a.H && j < a.W);
        const int pix = inside ? i * a.W + j : 0;
        T[k] = inside ? Tfin[pix] : 1.f;
        lst[k] = inside ? last[pix] : -1;
        gr[k] = inside ? gout[3 * pix + 0] : 0.f;
        gg[k] = inside ? gout[3 * pix + 1] : 0.f;
        gb[k] = inside ? gout[3 * pix + 2] : 0.f;
        acc[k] = 0.f;
        my_last = max(my_last, lst[k]);
    }
    const int wave_last = wave_max(my_last);
    int block_last = wave_last;
    if constexpr (NW > 1) {
        if (tid == 0) s_last = -1;
#pragma unroll
        for (int f = 0; f < 9; ++f) sG[f][tid] = 0.f;
        __syncthreads();
        if (__lane_id() == 0) atomicMax(&s_last, wave_last);
        __syncthreads();
        block_last = s_last;
    }
    float *const dst[9] = {grad.pos_x,  grad.pos_y,   grad.cov_xx,  grad.cov_xy, grad.cov_yy,
                           grad.color_r, grad.color_g, grad.color_b, grad.opacity};
    const int end = min(range.y, block_last + 1);
    for (int bend = end - 1; bend >= range.x; bend -= NT) {
        const int bs = min(NT, bend + 1 - range.x);
        __syncthreads();
        const int li = bend - tid;
        if (li >= range.x) {
            const float4 *r = a.rec + 3 * (size_t)a.flat[li];
            sA[tid] = r[0];
            sB[tid] = r[1];
            sC[tid] = r[2];
        }
        __syncthreads();
        for (int t = max(0, bend - wave_last); t < bs; ++t) {
            const int idx = bend - t;
            const float4 A = sA[t], B = sB[t];
            float v[9];
#pragma unroll
            for (int f = 0; f < 9; ++f) v[f] = 0.f;
            bool anyv = false;
#pragma unroll
            for (int k = 0; k < P; ++k) {
                if (idx > lst[k]) continue;
                AlphaEval e;
                if (!eval_alpha(A, B, px, py[k], e)) continue;
                anyv = true;
                const float4 C = sC[t];
                const float ra = __builtin_amdgcn_rcpf(fmaxf(1.f - e.alpha, 1e-6f));
                const float Tb = T[k] * ra;  // transmittance in front of this splat
                const float fac = e.alpha * Tb;
                v[5] = fmaf(fac, gr[k], v[5]);
                v[6] = fmaf(fac, gg[k], v[6]);
                v[7] = fmaf(fac, gb[k], v[7]);
                const float cg = fmaf(C.x, gr[k], fmaf(C.y, gg[k], C.z * gb[k]));
                float v_alpha;
                if constexpr (GRAD_MODE == DS_GRAD_REFERENCE)
                    v_alpha = Tb * (cg - acc[k]);
                else
                    v_alpha = Tb * cg - ra * acc[k];
                if (e.og <= kMaxAlpha) {
                    v[8] = fmaf(e.g, v_alpha, v[8]);
                    const float s = -0.5f * e.og * v_alpha;  // dL/d dist2
                    const float p = s * e.u, q = s * e.w;
                    v[0] += p;                   // x2 after the reduction
                    v[1] += q;                   // x2 after the reduction
                    v[2] = fmaf(-p, e.u, v[2]);
                    v[3] = fmaf(-p, e.w, v[3]);  // x2 after the reduction
                    v[4] = fmaf(-q, e.w, v[4]);
                }
                acc[k] = fmaf(fac, cg, acc[k]);
                T[k] = Tb;
            }
            if (__ballot(anyv) == 0) continue;
#pragma unroll
            for (int f = 0; f < 9; ++f) v[f] = wave_sum_to_63(v[f]);
            v[0] *= 2.f;
            v[1] *= 2.f;
            v[3] *= 2.f;
            if (__lane_id() == kWave - 1) {
                if constexpr (NW == 1) {
                    const uint32_t g = __float_as_uint(B.w);
#pragma unroll
                    for (int f = 0; f < 9; ++f)
                        if (v[f] != 0.f) atomicAdd(dst[f] + g, v[f]);
                } else {
#pragma unroll
                    for (int f = 0; f < 9; ++f) atomicAdd(&sG[f][t], v[f]);
                }
            }
        }
        if constexpr (NW > 1) {
            __syncthreads();
            if (tid < bs) {
                const uint32_t g = __float_as_uint(sB[tid].w);
#pragma unroll
                for (int f = 0; f < 9; ++f) {
                    const float val = sG[f][tid];
                    if (val != 0.f) {
                        atomicAdd(dst[f] + g, val);
                        sG[f][tid] = 0.f;
                    }
                }
            }
        }
    }
}

}  // namespace

#define DS_PIX_DISPATCH(P, ...)                               \
    switch (P) {                                              \
        case 1: { constexpr int PP = 1; __VA_ARGS__; } break; \
        case 2: { constexpr int PP = 2; __VA_ARGS__; } break; \
        default: { constexpr int PP = 4; __VA_ARGS__; } break; \
    }

ds_status launch_composite_fwd(const CompositeArgs &a, int pix_per_thread, float *out_rgb,
                               float *out_T, int32_t *out_last, const float *target,
                               float *loss, float *grad, hipStream_t s) {
    if (a.ntiles <= 0) return DS_OK;
    const float inv_npix = 1.f / ((float)a.W * (float)a.H);
    const int P = pix_per_thread;
    if (target) {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_fwd<true, PP>), dim3(a.ntiles),
                                              dim3(kTilePix / PP), 0, s, a, out_rgb, out_T,
                                              out_last, target, loss, grad, inv_npix));
    } else {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_fwd<false, PP>), dim3(a.ntiles),
                                              dim3(kTilePix / PP), 0, s, a, out_rgb, out_T,
                                              out_last, (const float *)nullptr, (float *)nullptr,
                                              (float *)nullptr, inv_npix));
    }
    DS_CHECK_LAUNCH("composite_fwd");
    return DS_OK;
}

ds_status launch_composite_bwd(const CompositeArgs &a, int pix_per_thread, int grad_mode,
                               const float *T, const int32_t *last, const float *grad_rgb,
                               const ds_splat2d &grad, hipStream_t s) {
    if (a.ntiles <= 0) return DS_OK;
    const int P = pix_per_thread;
    if (grad_mode == DS_GRAD_REFERENCE) {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_bwd<DS_GRAD_REFERENCE, PP>),
                                              dim3(a.ntiles), dim3(kTilePix / PP), 0, s, a, T,
                                              last, grad_rgb, grad));
    } else {
        DS_PIX_DISPATCH(P, hipLaunchKernelGGL((k_composite_bwd<DS_GRAD_EXACT, PP>),
                                              dim3(a.ntiles), dim3(kTilePix / PP), 0, s, a, T,
                                              last, grad_rgb, grad));
    }
    DS_CHECK_LAUNCH("composite_bwd");
    return DS_OK;
}

}  // namespace ds
