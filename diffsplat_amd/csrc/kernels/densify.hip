// Adaptive density control: clone, split and prune of splats between fit
// steps, after Kerbl et al., "3D Gaussian Splatting for Real-Time Radiance
// Field Rendering", SIGGRAPH 2023, sec. 5.2. The reference lists it as a TODO
// (README.md:10) and has no implementation; the rules below are the paper's
// (and its published training loop's densify_and_clone / densify_and_split /
// prune steps):
//   - every fit step adds, for each splat visible in that step, the length of
//     the loss gradient of its projected (NDC) centre to an accumulator and 1
//     to a visit count;
//   - a splat whose mean gradient length reaches grad_threshold is cloned
//     when its largest scale is at most clone_max_scale (a copy with the same
//     parameters is appended), and otherwise split: it is replaced by two
//     splats whose centres are drawn from the splat's own Gaussian
//     N(mu, R S S R^T) and whose scales are divided by 1.6;
//   - a splat (or a new splat) whose opacity is below prune_min_opacity, or
//     (when prune_max_scale > 0) whose largest scale exceeds prune_max_scale,
//     is removed.
// The output keeps the surviving splats in their order, then the clones,
// then the split samples (the paper's code appends new splats). Survivors
// keep their Adam moments; new splats start with zero moments.
//
// Everything is decided on the raw parameters, so the decisions are exact
// comparisons: the scale test compares the largest log-scale with
// log(max scale), the opacity test the logit with logit(min opacity).
// Split samples use a counter-based hash of (seed, splat, sample, draw),
// identical on the host oracle, and the Box-Muller transform.
#include "../ds_device.hpp"
#include "../ds_internal.hpp"
#include "sort_scan.hpp"

namespace ds {
namespace {

constexpr int kBlock = 256;

// splitmix64 finaliser
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Uniform in (0, 1]: 24 bits of the hash of (seed, splat i, sample s, draw k).
__device__ __forceinline__ float uniform01(uint64_t seed, uint32_t i, uint32_t s, uint32_t k) {
    const uint64_t h = mix64(seed ^ mix64(((uint64_t)i << 8) | (s << 2) | k));
    return (float)((uint32_t)(h >> 40) + 1u) * (1.f / 16777216.f);
}

__global__ __launch_bounds__(kBlock) void k_densify_accum(const float *__restrict__ gx,
                                                          const float *__restrict__ gy,
                                                          const uint32_t *__restrict__ ntiles,
                                                          float *__restrict__ accum,
                                                          uint32_t *__restrict__ visits, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || ntiles[i] == 0) return;
    const float x = gx[i], y = gy[i];
    accum[i] += sqrtf(x * x + y * y);
    visits[i] += 1u;
}

// Per splat: its action, whether it survives (0/1) and how many new splats it
// adds (0, 1 or 2).
__global__ __launch_bounds__(kBlock) void k_densify_classify(
    ds_gauss3d p, const float *__restrict__ accum, const uint32_t *__restrict__ visits,
    DensifyParams c, uint8_t *__restrict__ action, uint32_t *__restrict__ keep,
    uint32_t *__restrict__ added, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nv = visits[i];
    const float g = nv ? accum[i] / (float)nv : 0.f;
    const float ls = fmaxf(fmaxf(p.scale_x[i], p.scale_y[i]), p.scale_z[i]);
    const bool faint = p.logit_opacity[i] < c.min_logit;
    const bool hot = g >= c.grad_threshold;
    const bool small = ls <= c.clone_max_log_scale;
    const bool too_big = c.prune_big && ls > c.prune_max_log_scale;
    const bool child_too_big = c.prune_big && ls - c.split_log_div > c.prune_max_log_scale;
    int a;
    uint32_t k = 0, m = 0;
    if (faint) {
        a = DS_DENSIFY_PRUNE;
    } else if (hot && small) {
        a = DS_DENSIFY_CLONE;
        k = m = too_big ? 0u : 1u;  // the copy has the original's scale
    } else if (hot) {
        a = DS_DENSIFY_SPLIT;
        m = child_too_big ? 0u : 2u;
    } else {
        a = too_big ? DS_DENSIFY_PRUNE : DS_DENSIFY_KEEP;
        k = too_big ? 0u : 1u;
    }
    action[i] = (uint8_t)a;
    keep[i] = k;
    added[i] = m;
}

// Writes the new slab (field stride n_out = keeps + adds, read on the device)
// and the new Adam moments (the same field-major layout, any field order).
__global__ __launch_bounds__(kBlock) void k_densify_scatter(
    const float *__restrict__ src, const float *__restrict__ src_m,
    const float *__restrict__ src_v, int n, int nfields, const uint8_t *__restrict__ action,
    const uint32_t *__restrict__ keep, const uint32_t *__restrict__ added,
    const uint32_t *__restrict__ keep_off, const uint32_t *__restrict__ add_off,
    const uint32_t *__restrict__ totals, DensifyParams c, float *__restrict__ dst,
    float *__restrict__ dst_m, float *__restrict__ dst_v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t nkeep = totals[0];
    const size_t n_out = (size_t)nkeep + totals[1];
    const int a = action[i];
    const uint32_t nadd = added[i];
    if (keep[i]) {
        const size_t j = keep_off[i];
        for (int f = 0; f < nfields; ++f) {
            dst[f * n_out + j] = src[(size_t)f * n + i];
            dst_m[f * n_out + j] = src_m[(size_t)f * n + i];
            dst_v[f * n_out + j] = src_v[(size_t)f * n + i];
        }
    }
    if (nadd == 0) return;
    const size_t j0 = nkeep + add_off[i];
    for (uint32_t s = 0; s < nadd; ++s) {
        const size_t j = j0 + s;
        for (int f = 0; f < nfields; ++f) {
            dst[f * n_out + j] = src[(size_t)f * n + i];
            dst_m[f * n_out + j] = 0.f;
            dst_v[f * n_out + j] = 0.f;
        }
        if (a != DS_DENSIFY_SPLIT) continue;
        // a sample of N(mu, R S S R^T): mu + R (s * z), z ~ N(0, I)
        const float u0 = uniform01(c.seed, (uint32_t)i, s, 0);
        const float u1 = uniform01(c.seed, (uint32_t)i, s, 1);
        const float u2 = uniform01(c.seed, (uint32_t)i, s, 2);
        const float u3 = uniform01(c.seed, (uint32_t)i, s, 3);
        const float r0 = sqrtf(-2.f * logf(u0)), r1 = sqrtf(-2.f * logf(u2));
        const float t0 = 6.283185307179586f * u1, t1 = 6.283185307179586f * u3;
        const float zx = r0 * cosf(t0), zy = r0 * sinf(t0), zz = r1 * cosf(t1);
        const float sx = expf(src[3 * (size_t)n + i]) * zx;
        const float sy = expf(src[4 * (size_t)n + i]) * zy;
        const float sz = expf(src[5 * (size_t)n + i]) * zz;
        const Mat3 R = quat_to_rot(quat_normalize(src[6 * (size_t)n + i], src[7 * (size_t)n + i],
                                                  src[8 * (size_t)n + i], src[9 * (size_t)n + i]));
        const float d[3] = {R.r[0][0] * sx + R.r[0][1] * sy + R.r[0][2] * sz,
                            R.r[1][0] * sx + R.r[1][1] * sy + R.r[1][2] * sz,
                            R.r[2][0] * sx + R.r[2][1] * sy + R.r[2][2] * sz};
        for (int f = 0; f < 3; ++f) dst[f * n_out + j] = src[(size_t)f * n + i] + d[f];
        for (int f = 3; f < 6; ++f) dst[f * n_out + j] = src[(size_t)f * n + i] - c.split_log_div;
    }
}

}  // namespace

ds_status launch_densify_accum(const float *grad_x, const float *grad_y, const uint32_t *ntiles,
                               float *accum, uint32_t *visits, int n, hipStream_t s) {
    if (n <= 0) return DS_OK;
    hipLaunchKernelGGL(k_densify_accum, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       grad_x, grad_y, ntiles, accum, visits, n);
    DS_CHECK_LAUNCH("densify_accum");
    return DS_OK;
}

ds_status launch_densify(const ds_gauss3d &p, const float *m, const float *v, int nfields,
                         const float *accum, const uint32_t *visits, const DensifyParams &c,
                         const DensifyScratch &scr, float *dst, float *dst_m, float *dst_v,
                         hipStream_t s) {
    const int n = p.count;
    if (n <= 0) return DS_OK;
    const dim3 grid((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_densify_classify, grid, dim3(kBlock), 0, s, p, accum, visits, c,
                       scr.action, scr.keep, scr.added, n);
    DS_CHECK_LAUNCH("densify_classify");
    DS_TRY(fill_u32(scr.count, (uint32_t)n, 1, s));
    DS_TRY(scan_exclusive_u32(scr.keep, nullptr, scr.keep_off, ScanLen{scr.count, 1, 1},
                              (uint32_t)n, scr.part, scr.totals, nullptr, 0, s));
    DS_TRY(scan_exclusive_u32(scr.added, nullptr, scr.add_off, ScanLen{scr.count, 1, 1},
                              (uint32_t)n, scr.part, scr.totals + 1, nullptr, 0, s));
    hipLaunchKernelGGL(k_densify_scatter, grid, dim3(kBlock), 0, s, p.pos_x, m, v, n, nfields,
                       scr.action, scr.keep, scr.added, scr.keep_off, scr.add_off, scr.totals, c,
                       dst, dst_m, dst_v);
    DS_CHECK_LAUNCH("densify_scatter");
    return DS_OK;
}

}  // namespace ds
