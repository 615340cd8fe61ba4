// Kernels of the learned initializer (kernels in initnet.hip).
#pragma once

#include "../ds_device.hpp"
#include "../ds_internal.hpp"

namespace ds {

constexpr int kHeadChannels = 16;  // per-cell head values (9 used, padded for the MFMA tile)
constexpr float kSHC0 = 0.28209479177387814f;

// One convolution layer: NHWC input Hi x Wi x cin -> Ho x Wo x cout, square
// kernel ks (3 or 1, padding ks / 2), stride 1 or 2. kpad = ks*ks*cin and
// kpad2 = ks*ks*cout rounded up to 32 (the MFMA K step).
struct ConvShape {
    int cin, cout, ks, stride;
    int Hi, Wi, Ho, Wo;
    int kpad, kpad2;
};

// Geometry of the head: a gw x gh grid of cells of `cell` pixels over a W x H
// image seen by the pinhole camera at the origin looking down -z. Each cell's
// splat sits on the plane at `depth`, up to max_off pixels from the cell
// centre; log-scales are log_scale0 +- max_log_scale (z: their mean +
// log_scale_z); rotation is about the view axis; opacity logit = h + bias.
struct HeadGeom {
    int W, H, gw, gh;
    float cell, max_off, depth, tan_half_fovy, aspect;
    float log_scale0, max_log_scale, log_scale_z, logit_bias;
};

ds_status conv_fwd(const ConvShape &c, const __bf16 *in, const __bf16 *w, const float *bias,
                   void *out, bool relu, bool out_f32, hipStream_t s);
// d_in = conv^T(d_out) masked by (y_in > 0), the ReLU of the layer below.
ds_status conv_bwd_data(const ConvShape &c, const __bf16 *dout, const __bf16 *wt,
                        const __bf16 *y_in, __bf16 *din, hipStream_t s);
// dw [cout][kpad] += ..., db [cout] += ... (fp32, zeroed by the caller).
ds_status conv_bwd_weight(const ConvShape &c, const __bf16 *dout, const __bf16 *in, float *dw,
                          float *db, hipStream_t s);
// fp32 master weights -> bf16 [cout][kpad] and (if wt) bf16 [cin][kpad2].
ds_status weights_to_bf16(const ConvShape &c, const float *w, __bf16 *wb, __bf16 *wt,
                          hipStream_t s);
ds_status image_to_bf16(const float *rgb, int n, __bf16 *out, hipStream_t s);
ds_status to_bf16(const float *x, int n, __bf16 *out, hipStream_t s);
// head values [gw*gh][kHeadChannels] -> SH-0 slab (14 x gw*gh floats).
ds_status head_fwd(const float *h, const HeadGeom &g, float *slab, hipStream_t s);
// slab gradients -> head-value gradients (fp32 copy optional, bf16 always).
ds_status head_bwd(const float *h, const HeadGeom &g, const float *grad_slab, float *dh32,
                   __bf16 *dh, hipStream_t s);

}  // namespace ds
