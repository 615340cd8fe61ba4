// Device-side helpers shared by every kernel of the engine (gfx950, wave64).
//
// Math follows the reference's device helpers:
//   normalizeQuat / quatToRotMat / viewDir   src/cuda/math_ops.cuh:34-101
//   SH constants and band order              src/utils/sh_consts.h:7-35,
//                                            src/layers/gauss_activ_layer.cu:117-160
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ds_api.h"

namespace ds {

constexpr int kWave = 64;

// ---------------------------------------------------------------- SH constants
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f;
constexpr float SH_C2_1 = -1.0925484305920792f;
constexpr float SH_C2_2 = 0.3153915652525200f;
constexpr float SH_C2_3 = -1.0925484305920792f;
constexpr float SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f;
constexpr float SH_C3_1 = 2.8906114426405540f;
constexpr float SH_C3_2 = -0.4570457994644658f;
constexpr float SH_C3_3 = 0.3731763325901154f;
constexpr float SH_C3_4 = -0.4570457994644658f;
constexpr float SH_C3_5 = 1.4453057213202770f;
constexpr float SH_C3_6 = -0.5900435899266435f;

__host__ __device__ constexpr int sh_bands(int degree) {
    return degree <= 0 ? 0 : degree == 1 ? 3 : degree == 2 ? 8 : 15;
}

// ---------------------------------------------------------------- rotation
struct Quat {
    float w, x, y, z;
    float inv_norm;  // 1 / |q_raw|
};

__device__ __forceinline__ Quat quat_normalize(float w, float x, float y, float z) {
    float inv = rsqrtf(fmaxf(w * w + x * x + y * y + z * z, 1e-12f));
    return {w * inv, x * inv, y * inv, z * inv, inv};
}

// Row-major 3x3: r[row][col].
struct Mat3 {
    float r[3][3];
};

__device__ __forceinline__ Mat3 quat_to_rot(const Quat &q) {
    Mat3 m;
    const float w = q.w, x = q.x, y = q.y, z = q.z;
    m.r[0][0] = 1.f - 2.f * (y * y + z * z);
    m.r[0][1] = 2.f * (x * y - w * z);
    m.r[0][2] = 2.f * (x * z + w * y);
    m.r[1][0] = 2.f * (x * y + w * z);
    m.r[1][1] = 1.f - 2.f * (x * x + z * z);
    m.r[1][2] = 2.f * (y * z - w * x);
    m.r[2][0] = 2.f * (x * z - w * y);
    m.r[2][1] = 2.f * (y * z + w * x);
    m.r[2][2] = 1.f - 2.f * (x * x + y * y);
    return m;
}

// Unit vector from the splat toward the camera; |d|^2 floored at 1e-12.
__device__ __forceinline__ float3 view_dir(float cx, float cy, float cz, float px, float py,
                                           float pz) {
    float dx = cx - px, dy = cy - py, dz = cz - pz;
    float inv = rsqrtf(fmaxf(dx * dx + dy * dy + dz * dz, 1e-12f));
    return make_float3(dx * inv, dy * inv, dz * inv);
}

// Real SH basis (3DGS sign convention) for the 15 non-DC bands, in the
// reference's storage order. Only the first sh_bands(DEG) entries are set.
template <int DEG>
__device__ __forceinline__ void sh_basis_rest(float3 v, float *Y) {
    if constexpr (DEG >= 1) {
        Y[0] = -SH_C1 * v.y;
        Y[1] = SH_C1 * v.z;
        Y[2] = -SH_C1 * v.x;
    }
    if constexpr (DEG >= 2) {
        const float xx = v.x * v.x, yy = v.y * v.y, zz = v.z * v.z;
        const float xy = v.x * v.y, yz = v.y * v.z, xz = v.x * v.z;
        Y[3] = SH_C2_0 * xy;
        Y[4] = SH_C2_1 * yz;
        Y[5] = SH_C2_2 * (2.f * zz - xx - yy);
        Y[6] = SH_C2_3 * xz;
        Y[7] = SH_C2_4 * (xx - yy);
        if constexpr (DEG >= 3) {
            Y[8] = SH_C3_0 * v.y * (3.f * xx - yy);
            Y[9] = SH_C3_1 * xy * v.z;
            Y[10] = SH_C3_2 * v.y * (4.f * zz - xx - yy);
            Y[11] = SH_C3_3 * v.z * (2.f * zz - 3.f * xx - 3.f * yy);
            Y[12] = SH_C3_4 * v.x * (4.f * zz - xx - yy);
            Y[13] = SH_C3_5 * v.z * (xx - yy);
            Y[14] = SH_C3_6 * v.x * (xx - 3.f * yy);
        }
    }
}

// ---------------------------------------------------------------- wave64
// One Adam element (adamKernel, src/optimizers/adam.cu:12-37) given the
// step's learning rate and bias corrections bc1 = 1 / (1 - beta1^t), bc2 =
// 1 / (1 - beta2^t); shared by the optimiser kernel and the fused projection
// backward so that both compute the same expression.
__device__ __forceinline__ void adam_elem(float &p, float &m, float &v, float g, float lr,
                                          float bc1, float bc2, float beta1, float beta2,
                                          float eps) {
    m = beta1 * m + (1.f - beta1) * g;
    v = beta2 * v + (1.f - beta2) * g * g;
    p -= lr * (m * bc1) / (sqrtf(v * bc2) + eps);
}

// Learning rate of a group at step t: lr * decay^(t - 1) (decay 1: constant).
__device__ __forceinline__ float adam_lr(float lr, float decay, int32_t t) {
    return decay == 1.f ? lr : lr * powf(decay, (float)(t - 1));
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
    return v;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Grid-stride helpers for kernels whose element count lives in device memory.
__device__ __forceinline__ int64_t gtid() {
    return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }

}  // namespace ds
