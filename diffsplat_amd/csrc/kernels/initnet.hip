// Convolution kernels of the learned initializer (a CNN of six 3x3
// convolutions that maps a target image to a grid of 16x16-pixel cells with
// 16 splats each), on bf16 MFMA with fp32 accumulation: forward,
// backward-data and backward-weight as implicit GEMMs, plus the head that
// turns the CNN's per-cell outputs into splat parameters, its backward, its
// inverse (the regression targets of fitted splats) and the regression loss.
// No counterpart in the reference, which starts every fit from random splats
// (src/utils/splat_utils.h:10-51).
//
// Layout: activations NHWC bf16; weights [Cout][Kpad] bf16 with the GEMM
// index k = (kh * KS + kw) * Cin + ci, zero-padded to Kpad (a multiple of
// 32), and a transposed copy [Cin][Kpad2] with k = (kh * KS + kw) * Cout + co
// for the backward-data GEMM.
//
// v_mfma_f32_16x16x32_bf16 lane layout (checked on gfx950): A (16 x 32):
// lane l holds row l % 16, k = 8 * (l / 16) + j; B (32 x 16): column l % 16,
// k = 8 * (l / 16) + j; D (16 x 16): lane l, register r holds row
// 4 * (l / 16) + r, column l % 16.
#include "initnet.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace ds {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.f;
    return z;
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) {
    return *reinterpret_cast<const bf16x8 *>(p);  // 16-byte aligned by construction
}

__device__ __forceinline__ floatx4 mfma(const bf16x8 &a, const bf16x8 &b, const floatx4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- forward

// One wave per 16 output pixels x 16*TN output channels; 4 waves per
// workgroup along the pixels. out = relu?(conv(in) + bias).
template <int CIN, int COUT, int KS, int S, int TN, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv_fwd(const __bf16 *__restrict__ in, int Hi, int Wi,
                                                  const __bf16 *__restrict__ w, int kpad,
                                                  const float *__restrict__ bias, int Ho, int Wo,
                                                  void *__restrict__ out) {
    constexpr int PAD = KS / 2;
    constexpr int K = KS * KS * CIN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = Ho * Wo;
    const int row = (blockIdx.x * 4 + wave) * 16 + (lane & 15);  // output pixel of the A fragment
    const int co0 = blockIdx.y * 16 * TN;
    const int oh = row / Wo, ow = row % Wo;
    const bool row_ok = row < M;
    floatx4 acc[TN];
#pragma unroll
    for (int t = 0; t < TN; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kpad; k0 += 32) {
        const int kb = k0 + 8 * (lane >> 4);
        bf16x8 a = zero8();
        if constexpr (CIN % 8 == 0) {
            const int tap = kb / CIN, ci = kb % CIN;
            if (row_ok && tap < KS * KS) {
                const int ih = oh * S - PAD + tap / KS, iw = ow * S - PAD + tap % KS;
                if (ih >= 0 && ih < Hi && iw >= 0 && iw < Wi)
                    a = load8(in + ((size_t)ih * Wi + iw) * CIN + ci);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = kb + j;
                if (row_ok && k < K) {
                    const int tap = k / CIN, ci = k % CIN;
                    const int ih = oh * S - PAD + tap / KS, iw = ow * S - PAD + tap % KS;
                    if (ih >= 0 && ih < Hi && iw >= 0 && iw < Wi)
                        a[j] = in[((size_t)ih * Wi + iw) * CIN + ci];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            const int co = co0 + 16 * t + (lane & 15);
            acc[t] = mfma(a, load8(w + (size_t)co * kpad + kb), acc[t]);
        }
    }
    const int prow = (blockIdx.x * 4 + wave) * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < TN; ++t) {
        const int co = co0 + 16 * t + (lane & 15);
        const float b = bias[co];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = prow + r;
            if (p >= M) continue;
            float v = acc[t][r] + b;
            if (RELU) v = fmaxf(v, 0.f);
            if constexpr (OUT_F32)
                reinterpret_cast<float *>(out)[(size_t)p * COUT + co] = v;
            else
                reinterpret_cast<__bf16 *>(out)[(size_t)p * COUT + co] = (__bf16)v;
        }
    }
}

// ---------------------------------------------------------- backward data

// d_in[q][ci] = sum over taps and co of d_out[o(q, tap)][co] * w[co][tap][ci],
// then times the ReLU mask of the layer below (y_in > 0) when MASK. One wave
// per 16 input pixels x 16*TN input channels.
template <int CIN, int COUT, int KS, int S, int TN, bool MASK>
__global__ __launch_bounds__(256) void k_conv_bwd_data(
    const __bf16 *__restrict__ dout, int Ho, int Wo, const __bf16 *__restrict__ wt, int kpad2,
    const __bf16 *__restrict__ y_in, int Hi, int Wi, __bf16 *__restrict__ din) {
    constexpr int PAD = KS / 2;
    static_assert(COUT % 8 == 0, "vector loads of d_out need Cout % 8 == 0");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = Hi * Wi;
    const int row = (blockIdx.x * 4 + wave) * 16 + (lane & 15);
    const int ci0 = blockIdx.y * 16 * TN;
    const int qh = row / Wi, qw = row % Wi;
    const bool row_ok = row < M;
    floatx4 acc[TN];
#pragma unroll
    for (int t = 0; t < TN; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kpad2; k0 += 32) {
        const int kb = k0 + 8 * (lane >> 4);
        const int tap = kb / COUT, co = kb % COUT;
        bf16x8 a = zero8();
        if (row_ok && tap < KS * KS) {
            const int nh = qh + PAD - tap / KS, nw = qw + PAD - tap % KS;
            if (nh >= 0 && nw >= 0 && nh % S == 0 && nw % S == 0) {
                const int oh = nh / S, ow = nw / S;
                if (oh < Ho && ow < Wo) a = load8(dout + ((size_t)oh * Wo + ow) * COUT + co);
            }
        }
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            const int ci = ci0 + 16 * t + (lane & 15);
            acc[t] = mfma(a, load8(wt + (size_t)ci * kpad2 + kb), acc[t]);
        }
    }
    const int prow = (blockIdx.x * 4 + wave) * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < TN; ++t) {
        const int ci = ci0 + 16 * t + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = prow + r;
            if (q >= M) continue;
            float v = acc[t][r];
            if (MASK && !((float)y_in[(size_t)q * CIN + ci] > 0.f)) v = 0.f;
            din[(size_t)q * CIN + ci] = (__bf16)v;
        }
    }
}

// -------------------------------------------------------- backward weight

// dw[co][k] += sum over output pixels p of d_out[p][co] * im2col(in)[p][k].
// One wave per 16 co x 16 k tile and a chunk of pixels; fp32 atomics into
// dw (zeroed before).
template <int CIN, int COUT, int KS, int S>
__global__ __launch_bounds__(64) void k_conv_bwd_weight(const __bf16 *__restrict__ dout,
                                                        int Ho, int Wo,
                                                        const __bf16 *__restrict__ in, int Hi,
                                                        int Wi, int p_chunk,
                                                        float *__restrict__ dw, int kpad) {
    constexpr int PAD = KS / 2;
    constexpr int K = KS * KS * CIN;
    const int lane = threadIdx.x;
    const int co = blockIdx.x * 16 + (lane & 15);  // A row
    const int kk = blockIdx.y * 16 + (lane & 15);  // B column
    const int P = Ho * Wo;
    const int pbeg = blockIdx.z * p_chunk, pend = min(P, pbeg + p_chunk);
    const int tap = kk / CIN, ci = kk % CIN;
    const int kh = tap / KS, kw = tap % KS;
    const bool k_ok = kk < K;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int p0 = pbeg; p0 < pend; p0 += 32) {
        bf16x8 a = zero8(), b = zero8();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = p0 + 8 * (lane >> 4) + j;
            if (p < pend) {
                a[j] = dout[(size_t)p * COUT + co];
                if (k_ok) {
                    const int oh = p / Wo, ow = p % Wo;
                    const int ih = oh * S - PAD + kh, iw = ow * S - PAD + kw;
                    if (ih >= 0 && ih < Hi && iw >= 0 && iw < Wi)
                        b[j] = in[((size_t)ih * Wi + iw) * CIN + ci];
                }
            }
        }
        acc = mfma(a, b, acc);
    }
    const int col = blockIdx.y * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int orow = blockIdx.x * 16 + 4 * (lane >> 4) + r;
        if (col < K && acc[r] != 0.f) atomicAdd(&dw[(size_t)orow * kpad + col], acc[r]);
    }
}

// db[co] += sum over pixels of d_out[p][co]; RPB pixel rows of COUT <= 256
// channels in flight per workgroup (threads past RPB * COUT idle).
template <int COUT>
__global__ __launch_bounds__(256) void k_bias_grad(const __bf16 *__restrict__ dout, int P,
                                                   float *__restrict__ db) {
    static_assert(COUT <= 256, "one thread per channel");
    constexpr int RPB = 256 / COUT;
    __shared__ float s[COUT];
    const int c = threadIdx.x % COUT, r = threadIdx.x / COUT;
    if (threadIdx.x < COUT) s[threadIdx.x] = 0.f;
    __syncthreads();
    if (r < RPB) {
        float sum = 0.f;
        for (int p = blockIdx.x * RPB + r; p < P; p += gridDim.x * RPB)
            sum += (float)dout[(size_t)p * COUT + c];
        atomicAdd(&s[c], sum);
    }
    __syncthreads();
    if (threadIdx.x < COUT) atomicAdd(&db[threadIdx.x], s[threadIdx.x]);
}

// fp32 master weights [Cout][kpad] -> bf16 [Cout][kpad] and the transposed
// bf16 copy [Cin][kpad2] (k2 = tap * Cout + co).
__global__ __launch_bounds__(256) void k_weights_bf16(const float *__restrict__ w, int cout,
                                                      int cin, int taps, int kpad,
                                                      __bf16 *__restrict__ wb,
                                                      __bf16 *__restrict__ wt, int kpad2) {
    const int n = cout * kpad;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int co = i / kpad, k = i % kpad;
        const __bf16 v = (__bf16)w[i];
        wb[i] = v;
        if (wt && k < taps * cin) {
            const int tap = k / cin, ci = k % cin;
            wt[(size_t)ci * kpad2 + tap * cout + co] = v;
        }
    }
}

// fp32 HWC image in [0, 1] -> bf16 NHWC centred at 0.
__global__ __launch_bounds__(256) void k_image_bf16(const float *__restrict__ rgb, int n,
                                                    __bf16 *__restrict__ out) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        out[i] = (__bf16)(rgb[i] - 0.5f);
}

// fp32 -> bf16 copy.
__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ x, int n,
                                                 __bf16 *__restrict__ out) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        out[i] = (__bf16)x[i];
}

// ------------------------------------------------------------------ head

__device__ __forceinline__ float sigmoidf(float x) { return 1.f / (1.f + expf(-x)); }

// Head values of splat (ax, ay): in its output cell, sub-position j.
__device__ __forceinline__ size_t head_base(const HeadGeom &g, int ax, int ay) {
    const int cell = (ay / kCellSub) * (g.gw / kCellSub) + ax / kCellSub;
    const int j = (ay % kCellSub) * kCellSub + ax % kCellSub;
    return (size_t)cell * kHeadChannels + (size_t)j * kHeadValues;
}

// Per splat c (row-major over the gw x gh anchors): its head values -> its
// SH-0 parameters in the slab (field-major, n = gw * gh splats). The splat
// sits on the ray through its image position (u, v) at depth d. Its
// footprint is the log of its 2x2 apparent covariance (its size at depth
// g.depth): mean log-scale m = log_scale0 + max_log_scale tanh(h2) and the
// anisotropy vector (h3, h4) = r (cos 2 theta, sin 2 theta), half the
// log-scale difference a = max_aniso tanh(sqrt(r^2 + eps^2) / max_aniso)
// and the major axis at angle theta: log-scales m +- a, both shifted by
// log(d / g.depth) so that d orders splats without changing footprints.
constexpr float kAnisoEps2 = 1e-8f;

__global__ __launch_bounds__(256) void k_head_fwd(const float *__restrict__ h, HeadGeom g,
                                                  float *__restrict__ slab) {
    const int n = g.gw * g.gh;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int cx = c % g.gw, cy = c / g.gw;
    const float *hc = h + head_base(g, cx, cy);
    const float u = g.cell * (cx + 0.5f) + g.max_off * tanhf(hc[0]);
    const float v = g.cell * (cy + 0.5f) + g.max_off * tanhf(hc[1]);
    const float xn = 2.f * u / g.W - 1.f, yn = 1.f - 2.f * v / g.H;
    const float d = g.depth + g.max_dz * tanhf(hc[9]);
    slab[0 * n + c] = xn * d * g.tan_half_fovy * g.aspect;
    slab[1 * n + c] = yn * d * g.tan_half_fovy;
    slab[2 * n + c] = -d;
    const float m = g.log_scale0 + g.max_log_scale * tanhf(hc[2]) + logf(d / g.depth);
    const float r = sqrtf(hc[3] * hc[3] + hc[4] * hc[4] + kAnisoEps2);
    const float a = g.max_aniso * tanhf(r / g.max_aniso);
    slab[3 * n + c] = m + a;
    slab[4 * n + c] = m - a;
    slab[5 * n + c] = m + g.log_scale_z;
    const float th = 0.25f * atan2f(hc[4], hc[3]);  // half of theta
    slab[6 * n + c] = cosf(th);
    slab[7 * n + c] = 0.f;
    slab[8 * n + c] = 0.f;
    slab[9 * n + c] = sinf(th);
#pragma unroll
    for (int k = 0; k < 3; ++k) slab[(10 + k) * n + c] = (sigmoidf(hc[5 + k]) - 0.5f) / kSHC0;
    slab[13 * n + c] = hc[8] + g.logit_bias;
}

// Chain rule of k_head_fwd: slab gradients -> head-value gradients (bf16,
// the backward GEMMs' input).
__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ h, HeadGeom g,
                                                  const float *__restrict__ gs,
                                                  float *__restrict__ dh32,
                                                  __bf16 *__restrict__ dh) {
    const int n = g.gw * g.gh;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int cx = c % g.gw, cy = c / g.gw;
    const size_t base = head_base(g, cx, cy);
    const float *hc = h + base;
    float dv[kHeadValues];
    const float t0 = tanhf(hc[0]), t1 = tanhf(hc[1]), t9 = tanhf(hc[9]);
    const float u = g.cell * (cx + 0.5f) + g.max_off * t0;
    const float v = g.cell * (cy + 0.5f) + g.max_off * t1;
    const float xn = 2.f * u / g.W - 1.f, yn = 1.f - 2.f * v / g.H;
    const float d = g.depth + g.max_dz * t9;
    // position: u -> x_ndc -> pos_x, v -> y_ndc -> pos_y, d -> all three
    const float gx = gs[0 * n + c] * g.tan_half_fovy * g.aspect;
    const float gy = gs[1 * n + c] * g.tan_half_fovy;
    dv[0] = gx * d * (2.f / g.W) * g.max_off * (1.f - t0 * t0);
    dv[1] = -gy * d * (2.f / g.H) * g.max_off * (1.f - t1 * t1);
    // footprint: m in all three log-scales, +-a in the in-plane two
    const float t2 = tanhf(hc[2]);
    const float gm = gs[3 * n + c] + gs[4 * n + c] + gs[5 * n + c];
    dv[2] = gm * g.max_log_scale * (1.f - t2 * t2);
    const float r2 = hc[3] * hc[3] + hc[4] * hc[4] + kAnisoEps2, r = sqrtf(r2);
    const float ta = tanhf(r / g.max_aniso);
    const float ga = (gs[3 * n + c] - gs[4 * n + c]) * (1.f - ta * ta) / r;  // dL/dr / r
    // theta/2 = atan2(h4, h3) / 4; the quaternion (cos, 0, 0, sin) of it
    const float th = 0.25f * atan2f(hc[4], hc[3]);
    const float gth = -sinf(th) * gs[6 * n + c] + cosf(th) * gs[9 * n + c];
    dv[3] = ga * hc[3] - 0.25f * gth * hc[4] / r2;
    dv[4] = ga * hc[4] + 0.25f * gth * hc[3] / r2;
    // depth: the position along the ray and log(d / depth) in every log-scale
    dv[9] = (gx * xn + gy * yn - gs[2 * n + c] + gm / d) * g.max_dz * (1.f - t9 * t9);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = sigmoidf(hc[5 + k]);
        dv[5 + k] = gs[(10 + k) * n + c] * s * (1.f - s) / kSHC0;
    }
    dv[8] = gs[13 * n + c];
#pragma unroll
    for (int k = 0; k < kHeadValues; ++k) {
        if (dh32) dh32[base + k] = dv[k];
        dh[base + k] = (__bf16)dv[k];
    }
}

// Regression targets: the head values whose splat has the fitted splat's
// footprint, each head function inverted (a tanh or sigmoid argument clamped
// to +-kHeadClip of its range first):
//   - image position (u, v) of the fitted centre, and its depth;
//   - the footprint: the fitted 3D covariance projected along the view rays
//     (K Sigma K^T with K = [[1, 0, x/d], [0, 1, y/d]], the pinhole camera's
//     projection Jacobian divided by its focal terms), rescaled to the head
//     depth g.depth; from its eigenvalues the mean log-scale and half their
//     log difference a, from its major axis (angle theta) the anisotropy
//     vector's direction (cos 2 theta, sin 2 theta). The head's splats are
//     flat (z log-scale well below the in-plane ones), so this footprint is
//     theirs up to that flat extent;
//   - colour and opacity logit.
constexpr float kHeadClip = 0.98f;

__device__ __forceinline__ float atanh_clamped(float t) {
    t = fminf(fmaxf(t, -kHeadClip), kHeadClip);
    return 0.5f * logf((1.f + t) / (1.f - t));
}

__global__ __launch_bounds__(256) void k_head_targets(const float *__restrict__ slab, HeadGeom g,
                                                      float *__restrict__ tgt) {
    const int n = g.gw * g.gh;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int cx = c % g.gw, cy = c / g.gw;
    float *t = tgt + head_base(g, cx, cy);
    const float x = slab[0 * n + c], y = slab[1 * n + c];
    const float d = fmaxf(-slab[2 * n + c], 1e-3f);
    const float xn = x / (d * g.tan_half_fovy * g.aspect);
    const float yn = y / (d * g.tan_half_fovy);
    const float u = 0.5f * (xn + 1.f) * g.W, v = 0.5f * (1.f - yn) * g.H;
    t[0] = atanh_clamped((u - g.cell * (cx + 0.5f)) / g.max_off);
    t[1] = atanh_clamped((v - g.cell * (cy + 0.5f)) / g.max_off);
    t[9] = atanh_clamped((d - g.depth) / g.max_dz);
    // Sigma = R diag(s^2) R^T of the normalised quaternion (w, x, y, z)
    float qw = slab[6 * n + c], qx = slab[7 * n + c], qy = slab[8 * n + c], qz = slab[9 * n + c];
    const float qn = rsqrtf(fmaxf(qw * qw + qx * qx + qy * qy + qz * qz, 1e-30f));
    qw *= qn;
    qx *= qn;
    qy *= qn;
    qz *= qn;
    const float R[3][3] = {
        {1.f - 2.f * (qy * qy + qz * qz), 2.f * (qx * qy - qw * qz), 2.f * (qx * qz + qw * qy)},
        {2.f * (qx * qy + qw * qz), 1.f - 2.f * (qx * qx + qz * qz), 2.f * (qy * qz - qw * qx)},
        {2.f * (qx * qz - qw * qy), 2.f * (qy * qz + qw * qx), 1.f - 2.f * (qx * qx + qy * qy)}};
    const float s2[3] = {expf(2.f * slab[3 * n + c]), expf(2.f * slab[4 * n + c]),
                         expf(2.f * slab[5 * n + c])};
    // rows of K R: footprint A = sum_k s2_k (KR)_{.k} (KR)_{.k}^T
    const float ax_ = x / d, ay_ = y / d;
    float a = 0.f, b = 0.f, e = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float p = R[0][k] + ax_ * R[2][k], q = R[1][k] + ay_ * R[2][k];
        a += s2[k] * p * p;
        b += s2[k] * p * q;
        e += s2[k] * q * q;
    }
    // at depth g.depth: scale by (g.depth / d)^2
    const float k2 = (g.depth / d) * (g.depth / d);
    const float tr = 0.5f * (a + e), half = sqrtf(0.25f * (a - e) * (a - e) + b * b);
    const float l1 = k2 * (tr + half), l2 = fmaxf(k2 * (tr - half), 1e-12f * l1);
    const float L1 = 0.5f * logf(l1), L2 = 0.5f * logf(l2);
    t[2] = atanh_clamped((0.5f * (L1 + L2) - g.log_scale0) / g.max_log_scale);
    const float r = g.max_aniso * atanh_clamped(0.5f * (L1 - L2) / g.max_aniso);
    const float c2 = half > 0.f ? 0.5f * (a - e) / half : 1.f, s2_ = half > 0.f ? b / half : 0.f;
    t[3] = r * c2;
    t[4] = r * s2_;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float col = fminf(fmaxf(slab[(10 + k) * n + c] * kSHC0 + 0.5f, 1.f - kHeadClip),
                                kHeadClip);
        t[5 + k] = logf(col / (1.f - col));
    }
    t[8] = slab[13 * n + c] - g.logit_bias;
}

// Weighted squared error of the head values, mean over the splats, and its
// gradient.
__global__ __launch_bounds__(256) void k_head_regression(const float *__restrict__ h,
                                                         const float *__restrict__ tgt,
                                                         HeadGeom g, HeadWeights w,
                                                         __bf16 *__restrict__ dh, float *loss) {
    __shared__ float red[256 / kWave];
    const int n = g.gw * g.gh;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const float inv_n = 1.f / (float)n;
    float l = 0.f;
    if (c < n) {
        const size_t base = head_base(g, c % g.gw, c / g.gw);
#pragma unroll
        for (int k = 0; k < kHeadValues; ++k) {
            const float d = h[base + k] - tgt[base + k];
            dh[base + k] = (__bf16)(2.f * w.w[k] * d * inv_n);
            l += w.w[k] * d * d;
        }
    }
    if (!loss) return;
    l = wave_sum(l * inv_n);
    if (__lane_id() == 0) red[threadIdx.x / kWave] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

int blocks_for(long n, int cap = 4096) {
    long b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

// Launchers of one layer shape. Output tiles are 16 x 16*TN with TN up to 4.
// Column tiles per wave: the largest of 4, 3, 2, 1 dividing the 16-wide tiles.
constexpr int tiles_per_wave(int tiles) {
    return tiles % 4 == 0 ? 4 : tiles % 3 == 0 ? 3 : tiles % 2 == 0 ? 2 : 1;
}

template <int CI, int CO, int KS, int S>
ds_status launch_fwd(const ConvShape &c, const __bf16 *in, const __bf16 *w, const float *bias,
                     void *out, bool relu, bool out_f32, hipStream_t s) {
    static_assert(CO % 16 == 0, "Cout must be a multiple of 16");
    constexpr int TN = tiles_per_wave(CO / 16);
    const dim3 grid((c.Ho * c.Wo + 63) / 64, CO / (16 * TN));
    if (relu && !out_f32)
        hipLaunchKernelGGL((k_conv_fwd<CI, CO, KS, S, TN, true, false>), grid, dim3(256), 0, s,
                           in, c.Hi, c.Wi, w, c.kpad, bias, c.Ho, c.Wo, out);
    else if (!relu && out_f32)
        hipLaunchKernelGGL((k_conv_fwd<CI, CO, KS, S, TN, false, true>), grid, dim3(256), 0, s,
                           in, c.Hi, c.Wi, w, c.kpad, bias, c.Ho, c.Wo, out);
    else
        return set_error(DS_ERR_INVALID_ARG, "conv_fwd: unsupported epilogue");
    DS_CHECK_LAUNCH("conv_fwd");
    return DS_OK;
}

template <int CI, int CO, int KS, int S>
ds_status launch_bwd_data(const ConvShape &c, const __bf16 *dout, const __bf16 *wt,
                          const __bf16 *y_in, __bf16 *din, hipStream_t s) {
    if constexpr (CI % 16 == 0) {
        constexpr int TN = tiles_per_wave(CI / 16);
        const dim3 grid((c.Hi * c.Wi + 63) / 64, CI / (16 * TN));
        hipLaunchKernelGGL((k_conv_bwd_data<CI, CO, KS, S, TN, true>), grid, dim3(256), 0, s,
                           dout, c.Ho, c.Wo, wt, c.kpad2, y_in, c.Hi, c.Wi, din);
        DS_CHECK_LAUNCH("conv_bwd_data");
        return DS_OK;
    } else {
        return set_error(DS_ERR_INVALID_ARG, "conv_bwd_data: Cin must be a multiple of 16");
    }
}

template <int CI, int CO, int KS, int S>
ds_status launch_bwd_weight(const ConvShape &c, const __bf16 *dout, const __bf16 *in,
                            float *dw, float *db, hipStream_t s) {
    constexpr int K = KS * KS * CI;
    constexpr int kPixChunk = 1024;
    const int P = c.Ho * c.Wo;
    const dim3 grid(CO / 16, (K + 15) / 16, (P + kPixChunk - 1) / kPixChunk);
    hipLaunchKernelGGL((k_conv_bwd_weight<CI, CO, KS, S>), grid, dim3(64), 0, s, dout, c.Ho,
                       c.Wo, in, c.Hi, c.Wi, kPixChunk, dw, c.kpad);
    DS_CHECK_LAUNCH("conv_bwd_weight");
    hipLaunchKernelGGL((k_bias_grad<CO>), dim3(blocks_for(P, 256)), dim3(256), 0, s, dout, P,
                       db);
    DS_CHECK_LAUNCH("bias_grad");
    return DS_OK;
}

// Calls F<CI, CO, KS, S>(args) for the layer shape of c (the six layers of
// the initializer).
template <template <int, int, int, int> class F, typename... A>
ds_status dispatch_layer(const ConvShape &c, const char *what, A... args) {
#define DS_LAYER(CI, CO, KS, S)                                                 \
    if (c.cin == CI && c.cout == CO && c.ks == KS && c.stride == S)             \
        return F<CI, CO, KS, S>::run(c, args...);
    DS_LAYER(3, 32, 3, 2)
    DS_LAYER(32, 32, 3, 2)
    DS_LAYER(32, 64, 3, 2)
    DS_LAYER(64, 64, 3, 2)
    DS_LAYER(64, 64, 3, 1)
    DS_LAYER(64, kHeadChannels, 3, 1)
#undef DS_LAYER
    return set_error(DS_ERR_INVALID_ARG, std::string(what) + ": unsupported layer shape");
}

template <int CI, int CO, int KS, int S>
struct FwdOp {
    template <typename... A>
    static ds_status run(const ConvShape &c, A... a) { return launch_fwd<CI, CO, KS, S>(c, a...); }
};
template <int CI, int CO, int KS, int S>
struct BwdDataOp {
    template <typename... A>
    static ds_status run(const ConvShape &c, A... a) {
        return launch_bwd_data<CI, CO, KS, S>(c, a...);
    }
};
template <int CI, int CO, int KS, int S>
struct BwdWeightOp {
    template <typename... A>
    static ds_status run(const ConvShape &c, A... a) {
        return launch_bwd_weight<CI, CO, KS, S>(c, a...);
    }
};

}  // namespace

// ---------------------------------------------------------------- launchers

ds_status conv_fwd(const ConvShape &c, const __bf16 *in, const __bf16 *w, const float *bias,
                   void *out, bool relu, bool out_f32, hipStream_t s) {
    return dispatch_layer<FwdOp>(c, "conv_fwd", in, w, bias, out, relu, out_f32, s);
}

ds_status conv_bwd_data(const ConvShape &c, const __bf16 *dout, const __bf16 *wt,
                        const __bf16 *y_in, __bf16 *din, hipStream_t s) {
    return dispatch_layer<BwdDataOp>(c, "conv_bwd_data", dout, wt, y_in, din, s);
}

ds_status conv_bwd_weight(const ConvShape &c, const __bf16 *dout, const __bf16 *in, float *dw,
                          float *db, hipStream_t s) {
    return dispatch_layer<BwdWeightOp>(c, "conv_bwd_weight", dout, in, dw, db, s);
}

ds_status weights_to_bf16(const ConvShape &c, const float *w, __bf16 *wb, __bf16 *wt,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks_for((long)c.cout * c.kpad)), dim3(256), 0, s, w,
                       c.cout, c.cin, c.ks * c.ks, c.kpad, wb, wt, c.kpad2);
    DS_CHECK_LAUNCH("weights_bf16");
    return DS_OK;
}

ds_status image_to_bf16(const float *rgb, int n, __bf16 *out, hipStream_t s) {
    hipLaunchKernelGGL(k_image_bf16, dim3(blocks_for(n)), dim3(256), 0, s, rgb, n, out);
    DS_CHECK_LAUNCH("image_bf16");
    return DS_OK;
}

ds_status to_bf16(const float *x, int n, __bf16 *out, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, s, x, n, out);
    DS_CHECK_LAUNCH("to_bf16");
    return DS_OK;
}

ds_status head_fwd(const float *h, const HeadGeom &g, float *slab, hipStream_t s) {
    hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for((long)g.gw * g.gh, 1 << 20)), dim3(256), 0, s,
                       h, g, slab);
    DS_CHECK_LAUNCH("head_fwd");
    return DS_OK;
}

ds_status head_bwd(const float *h, const HeadGeom &g, const float *grad_slab, float *dh32,
                   __bf16 *dh, hipStream_t s) {
    hipLaunchKernelGGL(k_head_bwd, dim3(blocks_for((long)g.gw * g.gh, 1 << 20)), dim3(256), 0, s,
                       h, g, grad_slab, dh32, dh);
    DS_CHECK_LAUNCH("head_bwd");
    return DS_OK;
}

ds_status head_targets(const float *slab, const HeadGeom &g, float *target, hipStream_t s) {
    hipLaunchKernelGGL(k_head_targets, dim3(blocks_for((long)g.gw * g.gh, 1 << 20)), dim3(256), 0,
                       s, slab, g, target);
    DS_CHECK_LAUNCH("head_targets");
    return DS_OK;
}

ds_status head_regression(const float *h, const float *target, const HeadGeom &g,
                          const HeadWeights &w, __bf16 *dh, float *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_head_regression, dim3(blocks_for((long)g.gw * g.gh, 1 << 20)), dim3(256),
                       0, s, h, target, g, w, dh, loss);
    DS_CHECK_LAUNCH("head_regression");
    return DS_OK;
}

}  // namespace ds

// ====================================================================== net

namespace {

constexpr int kLayers = 6;

// Deterministic normal samples (splitmix64 + Box-Muller) for the weight init.
struct InitRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

int round32(int k) { return (k + 31) / 32 * 32; }

}  // namespace

struct ds_initnet {
    int W = 0, H = 0, gw = 0, gh = 0;
    ds::HeadGeom geom{};
    ds::HeadWeights loss_w = ds::kDefaultHeadWeights;
    ds::ConvShape L[kLayers]{};
    size_t w_off[kLayers]{}, b_off[kLayers]{};  // into the parameter buffer
    size_t n_params = 0;                          // padded, fp32
    int64_t n_export = 0;                         // unpadded
    ds::DevBuf<float> params, grads, m, v;
    ds::DevBuf<int32_t> step;
    ds::DevBuf<uint16_t> wb[kLayers], wt[kLayers];  // bf16 copies
    ds::DevBuf<uint16_t> x0, y[kLayers - 1];        // image and hidden activations (bf16)
    ds::DevBuf<float> h;                            // head values
    ds::DevBuf<uint16_t> dh, dy[kLayers - 1];       // gradients (bf16)
    bool has_fwd = false;
};

namespace ds {
namespace {

__bf16 *bf(DevBuf<uint16_t> &b) { return reinterpret_cast<__bf16 *>(b.ptr); }

ds_status net_weights_bf16(ds_initnet *n, hipStream_t s) {
    for (int l = 0; l < kLayers; ++l)
        DS_TRY(weights_to_bf16(n->L[l], n->params.ptr + n->w_off[l], bf(n->wb[l]),
                               l > 0 ? bf(n->wt[l]) : nullptr, s));
    return DS_OK;
}

ds_status net_forward(ds_initnet *n, const float *rgb, float *slab, hipStream_t s) {
    DS_TRY(image_to_bf16(rgb, n->W * n->H * 3, bf(n->x0), s));
    const __bf16 *in = bf(n->x0);
    for (int l = 0; l < kLayers; ++l) {
        const float *bias = n->params.ptr + n->b_off[l];
        if (l < kLayers - 1) {
            DS_TRY(conv_fwd(n->L[l], in, bf(n->wb[l]), bias, n->y[l].ptr, true, false, s));
            in = bf(n->y[l]);
        } else {
            DS_TRY(conv_fwd(n->L[l], in, bf(n->wb[l]), bias, n->h.ptr, false, true, s));
        }
    }
    if (slab) DS_TRY(head_fwd(n->h.ptr, n->geom, slab, s));
    n->has_fwd = true;
    return DS_OK;
}

// Gradients of all weights and biases (into n->grads) from the head-value
// gradient n->dh of the last forward.
ds_status net_backward_from_head(ds_initnet *n, hipStream_t s) {
    DS_TRY(fill_u32(n->grads.ptr, 0u, n->n_params, s));
    const __bf16 *dout = bf(n->dh);
    for (int l = kLayers - 1; l >= 0; --l) {
        const __bf16 *in = l > 0 ? bf(n->y[l - 1]) : bf(n->x0);
        DS_TRY(conv_bwd_weight(n->L[l], dout, in, n->grads.ptr + n->w_off[l],
                               n->grads.ptr + n->b_off[l], s));
        if (l > 0) {
            DS_TRY(conv_bwd_data(n->L[l], dout, bf(n->wt[l]), bf(n->y[l - 1]), bf(n->dy[l - 1]),
                                 s));
            dout = bf(n->dy[l - 1]);
        }
    }
    return DS_OK;
}

// The same from the slab gradient of the last forward.
ds_status net_backward(ds_initnet *n, const float *grad_slab, hipStream_t s) {
    DS_TRY(head_bwd(n->h.ptr, n->geom, grad_slab, nullptr, bf(n->dh), s));
    return net_backward_from_head(n, s);
}

ds_status net_update(ds_initnet *n, float lr, hipStream_t s) {
    const ds_adam_group g{n->params.ptr, n->grads.ptr, n->m.ptr, n->v.ptr, (int64_t)n->n_params,
                          lr, 1.f};
    const ds_adam_hparams hp{0.9f, 0.999f, 1e-8f};
    DS_TRY(ds_adam_step_dev(&g, 1, &hp, n->step.ptr, s));
    return net_weights_bf16(n, s);
}

// Copies between the padded device layout and the exported one (per layer:
// W[co][kh][kw][ci], then b[co]).
void pack_export(const ds_initnet *n, const std::vector<float> &dev, float *out) {
    size_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        const ConvShape &c = n->L[l];
        const int K = c.ks * c.ks * c.cin;
        for (int co = 0; co < c.cout; ++co)
            for (int k = 0; k < K; ++k) out[o++] = dev[n->w_off[l] + (size_t)co * c.kpad + k];
        for (int co = 0; co < c.cout; ++co) out[o++] = dev[n->b_off[l] + co];
    }
}

void unpack_export(const ds_initnet *n, const float *in, std::vector<float> &dev) {
    std::fill(dev.begin(), dev.end(), 0.f);
    size_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        const ConvShape &c = n->L[l];
        const int K = c.ks * c.ks * c.cin;
        for (int co = 0; co < c.cout; ++co)
            for (int k = 0; k < K; ++k) dev[n->w_off[l] + (size_t)co * c.kpad + k] = in[o++];
        for (int co = 0; co < c.cout; ++co) dev[n->b_off[l] + co] = in[o++];
    }
}

}  // namespace
}  // namespace ds

using namespace ds;

extern "C" ds_status ds_initnet_create(int width, int height, float fovy, uint64_t seed,
                                       ds_initnet **out) {
    clear_error();
    DS_REQUIRE(out, "NULL argument");
    DS_REQUIRE(width >= 16 && height >= 16 && width % 16 == 0 && height % 16 == 0 &&
                   width <= 4096 && height <= 4096,
               "initializer image size must be multiples of 16 in [16, 4096]");
    DS_REQUIRE(fovy > 0.f && fovy < 3.1f, "fovy out of range");
    auto n = std::make_unique<ds_initnet>();
    n->W = width;
    n->H = height;
    n->gw = width / 4;  // splat anchors, 4 pixels apart
    n->gh = height / 4;
    // six 3x3 convolutions: four of stride 2 down to one output cell per
    // 16x16 pixels, then two at that resolution, the last with the head's
    // values of the cell's 16 splats
    const int cin[kLayers] = {3, 32, 32, 64, 64, 64};
    const int cout[kLayers] = {32, 32, 64, 64, 64, kHeadChannels};
    const int ks[kLayers] = {3, 3, 3, 3, 3, 3};
    const int st[kLayers] = {2, 2, 2, 2, 1, 1};
    int hi = height, wi = width;
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        ConvShape &c = n->L[l];
        c.cin = cin[l];
        c.cout = cout[l];
        c.ks = ks[l];
        c.stride = st[l];
        c.Hi = hi;
        c.Wi = wi;
        c.Ho = (hi + 2 * (ks[l] / 2) - ks[l]) / st[l] + 1;
        c.Wo = (wi + 2 * (ks[l] / 2) - ks[l]) / st[l] + 1;
        c.kpad = round32(ks[l] * ks[l] * cin[l]);
        c.kpad2 = round32(ks[l] * ks[l] * cout[l]);
        n->w_off[l] = off;
        off += (size_t)c.cout * c.kpad;
        n->b_off[l] = off;
        off += (size_t)c.cout;
        n->n_export += (int64_t)c.cout * (ks[l] * ks[l] * cin[l] + 1);
        hi = c.Ho;
        wi = c.Wo;
    }
    n->n_params = off;
    DS_REQUIRE(hi * kCellSub == n->gh && wi * kCellSub == n->gw,
               "initializer layer sizes do not reach the cell grid");
    const float th = std::tan(0.5f * fovy);
    const float depth = 4.f;
    const float px_world = 2.f * depth * th / (float)height;
    n->geom = HeadGeom{width, height, n->gw, n->gh, 4.f, 16.f, depth, 1.f, th,
                       (float)width / (float)height,
                       std::log(2.f * px_world), 2.f, 2.f, -3.f, 1.f};
    // weights: He-normal for the ReLU layers, small for the head; biases 0
    std::vector<float> host(n->n_params, 0.f);
    InitRng rng{seed * 0x2545F4914F6CDD1Dull + 1};
    for (int l = 0; l < kLayers; ++l) {
        const ConvShape &c = n->L[l];
        const int K = c.ks * c.ks * c.cin;
        const double sd = (l < kLayers - 1) ? std::sqrt(2.0 / K) : 0.01 * std::sqrt(1.0 / K);
        for (int co = 0; co < c.cout; ++co)
            for (int k = 0; k < K; ++k)
                host[n->w_off[l] + (size_t)co * c.kpad + k] = (float)(sd * rng.normal());
    }
    ds_status stt = DS_OK;
    auto tryall = [&](ds_status s) { if (stt == DS_OK) stt = s; };
    tryall(n->params.alloc(n->n_params));
    tryall(n->grads.alloc(n->n_params));
    tryall(n->m.alloc(n->n_params));
    tryall(n->v.alloc(n->n_params));
    tryall(n->step.alloc(1));
    for (int l = 0; l < kLayers && stt == DS_OK; ++l) {
        const ConvShape &c = n->L[l];
        tryall(n->wb[l].alloc((size_t)c.cout * c.kpad));
        if (l > 0) tryall(n->wt[l].alloc((size_t)c.cin * c.kpad2));
        if (l < kLayers - 1) {
            tryall(n->y[l].alloc((size_t)c.Ho * c.Wo * c.cout));
            tryall(n->dy[l].alloc((size_t)c.Ho * c.Wo * c.cout));
        }
    }
    tryall(n->x0.alloc((size_t)width * height * 3));
    tryall(n->h.alloc((size_t)n->gw * n->gh * kHeadValues));
    tryall(n->dh.alloc((size_t)n->gw * n->gh * kHeadValues));
    if (stt != DS_OK) return stt;
    // the transposed copies' padding is never written: zero it once
    for (int l = 1; l < kLayers; ++l)
        DS_TRY(fill_u32(n->wt[l].ptr, 0u, n->wt[l].n / 2, nullptr));
    DS_TRY_HIP(hipMemcpy(n->params.ptr, host.data(), sizeof(float) * host.size(),
                         hipMemcpyHostToDevice));
    DS_TRY(fill_u32(n->m.ptr, 0u, n->n_params, nullptr));
    DS_TRY(fill_u32(n->v.ptr, 0u, n->n_params, nullptr));
    DS_TRY(fill_u32(reinterpret_cast<uint32_t *>(n->step.ptr), 0u, 1, nullptr));
    DS_TRY(net_weights_bf16(n.get(), nullptr));
    DS_TRY_HIP(hipDeviceSynchronize());
    *out = n.release();
    return DS_OK;
}

extern "C" ds_status ds_initnet_destroy(ds_initnet *n) {
    clear_error();
    if (n) (void)hipDeviceSynchronize();
    delete n;
    return DS_OK;
}

extern "C" ds_status ds_initnet_info(ds_initnet *n, int64_t *num_params, int32_t *num_splats) {
    clear_error();
    DS_REQUIRE(n, "NULL initializer");
    if (num_params) *num_params = n->n_export;
    if (num_splats) *num_splats = n->gw * n->gh;
    return DS_OK;
}

extern "C" ds_status ds_initnet_get_params(ds_initnet *n, float *host) {
    clear_error();
    DS_REQUIRE(n && host, "NULL argument");
    std::vector<float> dev(n->n_params);
    DS_TRY_HIP(hipDeviceSynchronize());
    DS_TRY_HIP(hipMemcpy(dev.data(), n->params.ptr, sizeof(float) * dev.size(),
                         hipMemcpyDeviceToHost));
    pack_export(n, dev, host);
    return DS_OK;
}

extern "C" ds_status ds_initnet_set_params(ds_initnet *n, const float *host) {
    clear_error();
    DS_REQUIRE(n && host, "NULL argument");
    std::vector<float> dev(n->n_params);
    unpack_export(n, host, dev);
    DS_TRY_HIP(hipDeviceSynchronize());
    DS_TRY_HIP(hipMemcpy(n->params.ptr, dev.data(), sizeof(float) * dev.size(),
                         hipMemcpyHostToDevice));
    DS_TRY(fill_u32(n->m.ptr, 0u, n->n_params, nullptr));
    DS_TRY(fill_u32(n->v.ptr, 0u, n->n_params, nullptr));
    DS_TRY(fill_u32(reinterpret_cast<uint32_t *>(n->step.ptr), 0u, 1, nullptr));
    DS_TRY(net_weights_bf16(n, nullptr));
    DS_TRY_HIP(hipDeviceSynchronize());
    return DS_OK;
}

extern "C" ds_status ds_initnet_infer(ds_initnet *n, const float *dev_rgb, float *dev_slab,
                                      ds_stream_t s) {
    clear_error();
    DS_REQUIRE(n && dev_rgb && dev_slab, "NULL argument");
    return net_forward(n, dev_rgb, dev_slab, as_stream(s));
}

extern "C" ds_status ds_initnet_head(ds_initnet *n, float *host_h) {
    clear_error();
    DS_REQUIRE(n && host_h, "NULL argument");
    DS_REQUIRE(n->has_fwd, "no forward yet");
    DS_TRY_HIP(hipDeviceSynchronize());
    DS_TRY_HIP(hipMemcpy(host_h, n->h.ptr, sizeof(float) * n->h.n, hipMemcpyDeviceToHost));
    return DS_OK;
}

extern "C" ds_status ds_initnet_grads(ds_initnet *n, const float *dev_rgb,
                                      const float *dev_grad_slab, float *host_grads) {
    clear_error();
    DS_REQUIRE(n && dev_rgb && dev_grad_slab && host_grads, "NULL argument");
    DS_TRY(net_forward(n, dev_rgb, nullptr, nullptr));
    DS_TRY(net_backward(n, dev_grad_slab, nullptr));
    std::vector<float> dev(n->n_params);
    DS_TRY_HIP(hipDeviceSynchronize());
    DS_TRY_HIP(hipMemcpy(dev.data(), n->grads.ptr, sizeof(float) * dev.size(),
                         hipMemcpyDeviceToHost));
    pack_export(n, dev, host_grads);
    return DS_OK;
}

extern "C" ds_status ds_initnet_head_targets(ds_initnet *n, const float *dev_slab,
                                             float *dev_targets, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(n && dev_slab && dev_targets, "NULL argument");
    return head_targets(dev_slab, n->geom, dev_targets, as_stream(s));
}

extern "C" ds_status ds_initnet_set_loss_weights(ds_initnet *n, const float *weights) {
    clear_error();
    DS_REQUIRE(n, "NULL argument");
    if (!weights) {
        n->loss_w = kDefaultHeadWeights;
        return DS_OK;
    }
    for (int k = 0; k < kHeadValues; ++k)
        DS_REQUIRE(std::isfinite(weights[k]) && weights[k] >= 0.f,
                   "loss weights must be finite and non-negative");
    for (int k = 0; k < kHeadValues; ++k) n->loss_w.w[k] = weights[k];
    return DS_OK;
}

extern "C" ds_status ds_initnet_train_step_supervised(ds_initnet *n, const float *dev_rgb,
                                                      const float *dev_targets, float lr,
                                                      float *dev_loss, ds_stream_t st) {
    clear_error();
    DS_REQUIRE(n && dev_rgb && dev_targets, "NULL argument");
    hipStream_t s = as_stream(st);
    DS_TRY(net_forward(n, dev_rgb, nullptr, s));
    DS_TRY(head_regression(n->h.ptr, dev_targets, n->geom, n->loss_w, bf(n->dh), dev_loss, s));
    DS_TRY(net_backward_from_head(n, s));
    return net_update(n, lr, s);
}

extern "C" ds_status ds_initnet_train_step(ds_initnet *n, ds_fitter *f, const float *dev_rgb,
                                           float lr) {
    clear_error();
    DS_REQUIRE(n && f && dev_rgb, "NULL argument");
    ds_gauss3d g{};
    DS_TRY(ds_fitter_views(f, &g, nullptr, nullptr, nullptr));
    DS_REQUIRE(g.count == n->gw * n->gh && g.sh_num_bands == 0,
               "the fitter must hold one SH-0 splat per initializer cell");
    float *pslab = nullptr;
    const float *gslab = nullptr;
    DS_TRY(ds_fitter_param_slabs(f, &pslab, &gslab));
    ds_stream_t fs = nullptr;
    DS_TRY(ds_fitter_get_stream(f, &fs));
    hipStream_t s = as_stream(fs);
    DS_TRY(net_forward(n, dev_rgb, pslab, s));
    DS_TRY(ds_fitter_set_target_device(f, dev_rgb));
    DS_TRY(ds_fitter_forward_backward(f));
    DS_TRY(net_backward(n, gslab, s));
    return net_update(n, lr, s);
}
