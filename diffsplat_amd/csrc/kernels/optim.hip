// MSE loss forward/backward, the multi-tensor Adam step and the fill kernel.
//
// Reference semantics:
//   mseLossKernel / mseGradKernel   src/layers/mse_loss_layer.cu:22-67
//   adamKernel / Adam::step         src/optimizers/adam.cu:12-37, 59-73
#include "../ds_device.hpp"
#include "../ds_internal.hpp"

namespace ds {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * kBlock;
    // 4 words per thread per iteration where aligned.
    for (; i < n; i += stride) p[i] = v;
}

// loss += sum_px sum_c (p - t)^2 / npix; one atomic per block.
__global__ __launch_bounds__(kBlock) void k_mse_fwd(const float *__restrict__ p,
                                                    const float *__restrict__ t, int64_t npix,
                                                    float *loss) {
    __shared__ float wsum[kBlock / kWave];
    const float scale = 1.f / (float)npix;
    float acc = 0.f;
    for (int64_t i = gtid(); i < npix; i += gstride()) {
        const float dr = p[3 * i + 0] - t[3 * i + 0];
        const float dg = p[3 * i + 1] - t[3 * i + 1];
        const float db = p[3 * i + 2] - t[3 * i + 2];
        acc += (dr * dr + dg * dg + db * db) * scale;
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = 0.f;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) b += wsum[w];
        atomicAdd(loss, b);
    }
}

__global__ __launch_bounds__(kBlock) void k_mse_bwd(const float *__restrict__ p,
                                                    const float *__restrict__ t, int64_t n3,
                                                    float scale, float *__restrict__ g) {
    for (int64_t i = gtid(); i < n3; i += gstride()) g[i] = scale * (p[i] - t[i]);
}

struct AdamTable {
    float *param[DS_ADAM_MAX_GROUPS];
    const float *grad[DS_ADAM_MAX_GROUPS];
    float *m[DS_ADAM_MAX_GROUPS];
    float *v[DS_ADAM_MAX_GROUPS];
    int64_t n[DS_ADAM_MAX_GROUPS];
    float lr[DS_ADAM_MAX_GROUPS];
    float decay[DS_ADAM_MAX_GROUPS];  // lr * decay^(t-1); 1: constant
    int count;
};

__global__ void k_adam_tick(int32_t *step) { *step += 1; }

// All groups in one launch: grid row blockIdx.y updates group blockIdx.y,
// grid-striding inside it along x. The groups' memory round trips overlap
// (a thread that walked the groups in turn would wait on each: stores to one
// group may alias the next group's loads).
__global__ __launch_bounds__(kBlock) void k_adam(AdamTable tab, float beta1, float beta2,
                                                 float eps, int32_t step_host,
                                                 const int32_t *step_dev) {
    const int k = blockIdx.y;
    const int32_t t = step_dev ? *step_dev : step_host;
    const float bc1 = 1.f / (1.f - powf(beta1, (float)t));
    const float bc2 = 1.f / (1.f - powf(beta2, (float)t));
    float *__restrict__ P = tab.param[k];
    const float *__restrict__ G = tab.grad[k];
    float *__restrict__ M = tab.m[k];
    float *__restrict__ V = tab.v[k];
    const float lr = adam_lr(tab.lr[k], tab.decay[k], t);
    const int64_t n = tab.n[k];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float p = P[i], m = M[i], v = V[i];
        adam_elem(p, m, v, G[i], lr, bc1, bc2, beta1, beta2, eps);
        M[i] = m;
        V[i] = v;
        P[i] = p;
    }
}

int grid_cap(int64_t n, int cap) {
    int64_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

ds_status fill_u32(void *p, uint32_t v, size_t count, hipStream_t s) {
    if (count == 0) return DS_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_cap((int64_t)count, 4096)), dim3(kBlock), 0, s,
                       (uint32_t *)p, v, count);
    DS_CHECK_LAUNCH("fill_u32");
    return DS_OK;
}

ds_status launch_mse_fwd(const float *p, const float *t, int64_t npix, float *loss,
                         hipStream_t s) {
    DS_TRY(fill_u32(loss, 0u, 1, s));
    if (npix <= 0) return DS_OK;
    hipLaunchKernelGGL(k_mse_fwd, dim3(grid_cap(npix, 2048)), dim3(kBlock), 0, s, p, t, npix,
                       loss);
    DS_CHECK_LAUNCH("mse_fwd");
    return DS_OK;
}

ds_status launch_mse_bwd(const float *p, const float *t, int64_t npix, float *grad,
                         hipStream_t s) {
    if (npix <= 0) return DS_OK;
    hipLaunchKernelGGL(k_mse_bwd, dim3(grid_cap(3 * npix, 4096)), dim3(kBlock), 0, s, p, t,
                       3 * npix, 2.f / (float)npix, grad);
    DS_CHECK_LAUNCH("mse_bwd");
    return DS_OK;
}

ds_status launch_adam(const ds_adam_group *groups, int n_groups, const ds_adam_hparams &hp,
                      int32_t step, int32_t *step_dev, hipStream_t s, bool tick) {
    AdamTable tab{};
    int64_t total = 0, largest = 0;
    for (int k = 0; k < n_groups; ++k) {
        tab.param[k] = groups[k].param;
        tab.grad[k] = groups[k].grad;
        tab.m[k] = groups[k].m;
        tab.v[k] = groups[k].v;
        tab.n[k] = groups[k].n;
        tab.lr[k] = groups[k].lr;
        tab.decay[k] = (groups[k].decay > 0.f) ? groups[k].decay : 1.f;
        total += groups[k].n;
        if (groups[k].n > largest) largest = groups[k].n;
    }
    tab.count = n_groups;
    if (step_dev && tick) {
        hipLaunchKernelGGL(k_adam_tick, dim3(1), dim3(1), 0, s, step_dev);
        DS_CHECK_LAUNCH("adam_tick");
    }
    if (total == 0) return DS_OK;
    hipLaunchKernelGGL(k_adam, dim3(grid_cap(largest, 2048), n_groups), dim3(kBlock), 0, s, tab,
                       hp.beta1, hp.beta2, hp.epsilon, step, (const int32_t *)step_dev);
    DS_CHECK_LAUNCH("adam");
    return DS_OK;
}

}  // namespace ds
