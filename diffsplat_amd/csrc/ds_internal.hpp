// Host-side internals shared by the C-ABI implementation files: error
// reporting, launch-wrapper declarations and small RAII device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>
#include <utility>

#include "../../include/ds_api.h"

namespace ds {

// Records a detail message for ds_last_error() and returns `st`.
ds_status set_error(ds_status st, const std::string &msg);
void clear_error();

#define DS_TRY_HIP(expr)                                                                  \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::ds::set_error(_e == hipErrorOutOfMemory ? DS_ERR_OUT_OF_MEMORY        \
                                                             : DS_ERR_HIP,                \
                                   std::string(#expr) + ": " + hipGetErrorString(_e));    \
    } while (0)

#define DS_CHECK_LAUNCH(name)                                                             \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess)                                                             \
            return ::ds::set_error(DS_ERR_HIP,                                            \
                                   std::string("launch ") + (name) + ": " +               \
                                       hipGetErrorString(_e));                            \
    } while (0)

#define DS_TRY(expr)                                  \
    do {                                              \
        ds_status _s = (expr);                        \
        if (_s != DS_OK) return _s;                   \
    } while (0)

#define DS_REQUIRE(cond, msg)                                             \
    do {                                                                  \
        if (!(cond)) return ::ds::set_error(DS_ERR_INVALID_ARG, (msg));   \
    } while (0)

inline hipStream_t as_stream(ds_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Move-only device allocation.
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    ds_status alloc(size_t count) {
        if (count == n && ptr) return DS_OK;
        release();
        if (count == 0) return DS_OK;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            return set_error(DS_ERR_OUT_OF_MEMORY, "hipMalloc of " +
                                                       std::to_string(count * sizeof(T)) +
                                                       " bytes: " + hipGetErrorString(e));
        }
        n = count;
        return DS_OK;
    }
    operator T *() const { return ptr; }
};

// ------------------------------------------------------------- launchers
// Each returns DS_OK or the launch error; all are asynchronous on `s`.

ds_status fill_u32(void *p, uint32_t v, size_t count, hipStream_t s);

// activation / projection (kernels/activ_project.hip)
struct PV {
    float m[16];
};
struct Cam3 {
    float x, y, z;
};
ds_status launch_activ_fwd(const ds_gauss3d &in, const ds_splat3d &out, int deg, Cam3 cam,
                           hipStream_t s);
ds_status launch_activ_bwd(const ds_gauss3d &in, const ds_splat3d &out, const ds_splat3d &gout,
                           const ds_gauss3d &gin, int deg, Cam3 cam, int mode, hipStream_t s);
ds_status launch_project_fwd(const ds_splat3d &in, const ds_splat2d &out, const PV &pv,
                             hipStream_t s);
ds_status launch_project_bwd(const ds_splat3d &in, const ds_splat2d &gout,
                             const ds_splat3d &gin, const PV &pv, int mode, hipStream_t s);
ds_status launch_activ_project_fwd(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                   Cam3 cam, const PV &pv, hipStream_t s);
ds_status launch_activ_project_bwd(const ds_gauss3d &in, const ds_splat2d &gout,
                                   const ds_gauss3d &gin, int deg, Cam3 cam, const PV &pv,
                                   int mode, hipStream_t s);
// Many fits at once: splat i uses the camera of instance splat_inst[i] in
// cam_tab (kCamRecF4 float4 per instance: the 16 PV floats column-major, then
// the camera position in x, y, z).
constexpr int kCamRecF4 = 5;
ds_status launch_activ_project_fwd_multi(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                         const float4 *cam_tab, const uint16_t *splat_inst,
                                         hipStream_t s);
ds_status launch_activ_project_bwd_multi(const ds_gauss3d &in, const ds_splat2d &gout,
                                         const ds_gauss3d &gin, int deg, const float4 *cam_tab,
                                         const uint16_t *splat_inst, int mode, hipStream_t s);

// The fitter's Adam step fused into the activation/projection backward (no
// counterpart in the reference; the reference's ImageFitter runs Adam as its
// own pass): after a splat's gradients are written, each of its parameters
// takes its Adam step. Field f of the table is, in order: position x, y, z,
// log-scale x, y, z, rotation w, x, y, z, SH DC r, g, b, opacity logit (14),
// then the SH rest r, g, b slabs (band-major, n * bands each) when the degree
// is above 0. step: the device step counter, already incremented.
constexpr int kAdamFields = 17;
struct AdamFields {
    float *p[kAdamFields];
    float *m[kAdamFields];
    float *v[kAdamFields];
    float lr[kAdamFields];
    float decay[kAdamFields];
    float beta1, beta2, eps;
    const int32_t *step;
};
ds_status launch_activ_project_bwd_adam(const ds_gauss3d &in, const ds_splat2d &gout,
                                        const ds_gauss3d &gin, int deg, Cam3 cam, const PV &pv,
                                        int mode, const AdamFields &af, hipStream_t s);

// loss / optimiser (kernels/optim.hip)
ds_status launch_mse_fwd(const float *p, const float *t, int64_t npix, float *loss,
                         hipStream_t s);
ds_status launch_mse_bwd(const float *p, const float *t, int64_t npix, float *grad,
                         hipStream_t s);
// With step_dev, *step_dev is incremented first (tick) and used; tick =
// false uses it as it is (a kernel earlier in the step incremented it).
ds_status launch_adam(const ds_adam_group *groups, int n_groups, const ds_adam_hparams &hp,
                      int32_t step, int32_t *step_dev, hipStream_t s, bool tick = true);

// The work of a rasterizer forward's first kernel, in neutral terms so that
// another kernel of a fitter step can do it instead (no counterpart in the
// reference): word ranges set to zero (the word at set_word to set_value
// instead), a counter stepped by one and never to 0 (epoch), and a fitter
// step's prelude: the Splat2D gradient arrays zeroed (ngrad floats each) and
// the Adam step counter incremented (tick may be NULL).
constexpr int kStartRanges = 6;
struct ForwardStart {
    uint32_t *zero[kStartRanges];
    uint32_t nzero[kStartRanges];
    uint32_t *set_word;
    uint32_t set_value;
    uint32_t *epoch;
    int32_t *tick;
    float *grad[10];
    int ngrad;
};

// One fitter step's rasterizer work with its prelude folded in (no
// counterpart in the reference): the first kernel of the forward also zeroes
// the Splat2D gradients grad_in (ImageFitter's zeroGrad) and increments the
// Adam step counter *adam_tick (the optimizer then runs with tick = false);
// then the forward with the fused MSE and the backward, as ds_raster_fwd_bwd.
// With started, that first kernel's work (raster_step_start) was already
// done by an earlier kernel of the step and is not launched.
ds_status raster_step_fwd_bwd(ds_raster *r, const ds_splat2d &in, float *out_rgb,
                              const float *target, float *loss, float *grad_rgb,
                              const ds_splat2d &grad_in, int32_t *adam_tick, hipStream_t s,
                              bool started = false);
// Adaptive density control (kernels/densify.hip). Thresholds are on the raw
// parameters: log-scales and opacity logits.
struct DensifyParams {
    float grad_threshold;       // mean |dL/d(NDC centre)| over visited steps
    float clone_max_log_scale;  // clone when the largest log-scale is <= this, else split
    float min_logit;            // prune when the opacity logit is below this
    float prune_max_log_scale;  // prune when the largest log-scale exceeds this (prune_big)
    float split_log_div;        // log of the split's scale divisor (1.6)
    int prune_big;
    uint64_t seed;              // split samples
};
// Device scratch of one densification of n splats (all n words unless
// noted): actions (bytes), survive flags, added counts, their exclusive
// scans, the two totals (survivors, added), the element count word and
// kSpineMax scan words.
struct DensifyScratch {
    uint8_t *action;
    uint32_t *keep, *added, *keep_off, *add_off;
    uint32_t *totals;
    uint32_t *count;
    uint32_t *part;
};
// Adds |dL/d(NDC centre)| and 1 visit for every splat with a tile (visible).
ds_status launch_densify_accum(const float *grad_x, const float *grad_y, const uint32_t *ntiles,
                               float *accum, uint32_t *visits, int n, hipStream_t s);
// Classifies, scans and scatters: dst, dst_m, dst_v receive nfields fields of
// stride n_out = totals[0] + totals[1] (at most 2 n) each; m and v are the
// Adam moments of p's slab in any field order.
ds_status launch_densify(const ds_gauss3d &p, const float *m, const float *v, int nfields,
                         const float *accum, const uint32_t *visits, const DensifyParams &c,
                         const DensifyScratch &scr, float *dst, float *dst_m, float *dst_v,
                         hipStream_t s);
// Per-splat tile counts of the rasterizer's last forward (0: culled).
const uint32_t *raster_splat_tiles(ds_raster *r);

// Reads the rasterizer's sticky look-back error flag (synchronises s); if a
// look-back gave up since the last read, clears the flag and returns
// DS_ERR_STATE with the detail set.
ds_status raster_take_error(ds_raster *r, hipStream_t s);
// The first kernel's work of raster_step_fwd_bwd for n splats.
ds_status raster_step_start(ds_raster *r, int n, float *loss, const ds_splat2d &grad_in,
                            int32_t *adam_tick, ForwardStart *out);
// Activation + projection that also does a forward's first-kernel work (n >
// 0: every thread of its grid takes a share).
ds_status launch_activ_project_fwd_start(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                         Cam3 cam, const PV &pv, const ForwardStart &fs,
                                         hipStream_t s);
// The same for many fits at once (cameras per instance, as
// launch_activ_project_fwd_multi), and the fused Adam backward likewise.
ds_status launch_activ_project_fwd_start_multi(const ds_gauss3d &in, const ds_splat2d &out,
                                               int deg, const float4 *cam_tab,
                                               const uint16_t *splat_inst,
                                               const ForwardStart &fs, hipStream_t s);
ds_status launch_activ_project_bwd_adam_multi(const ds_gauss3d &in, const ds_splat2d &gout,
                                              const ds_gauss3d &gin, int deg,
                                              const float4 *cam_tab, const uint16_t *splat_inst,
                                              int mode, const AdamFields &af, hipStream_t s);

}  // namespace ds
