// Kernels of the learned initializer (kernels in initnet.hip).
#pragma once

#include "../ds_device.hpp"
#include "../ds_internal.hpp"

namespace ds {

// per splat: offset x, y, mean log-scale, anisotropy (2), rgb, opacity, depth
constexpr int kHeadValues = 10;
constexpr int kCellSub = 4;        // splats per cell side (16 per cell)
constexpr int kHeadChannels = kHeadValues * kCellSub * kCellSub;  // per cell: 160
constexpr float kSHC0 = 0.28209479177387814f;

// One convolution layer: NHWC input Hi x Wi x cin -> Ho x Wo x cout, square
// kernel ks (3 or 1, padding ks / 2), stride 1 or 2. kpad = ks*ks*cin and
// kpad2 = ks*ks*cout rounded up to 32 (the MFMA K step).
struct ConvShape {
    int cin, cout, ks, stride;
    int Hi, Wi, Ho, Wo;
    int kpad, kpad2;
};

// Geometry of the head: a gw x gh grid of splat anchors `cell` pixels apart
// over a W x H image seen by the pinhole camera at the origin looking down
// -z; the network's output cells hold kCellSub x kCellSub anchors each
// (kHeadValues values per splat, kHeadChannels per cell). A splat's image
// position is up to max_off pixels from its anchor, at a depth within
// depth +- max_dz; its apparent log-scales (their size at `depth`) have
// their mean within log_scale0 +- max_log_scale and differ from it by up to
// max_aniso (z: the mean + log_scale_z, a flat splat); rotation is about the
// view axis; opacity logit = h + bias. Splat (ax, ay) is slab element
// ay * gw + ax.
struct HeadGeom {
    int W, H, gw, gh;
    float cell, max_off, depth, max_dz, tan_half_fovy, aspect;
    float log_scale0, max_log_scale, max_aniso, log_scale_z, logit_bias;
};

// Per head value, its weight in the supervised regression loss.
struct HeadWeights {
    float w[kHeadValues];
};
constexpr HeadWeights kDefaultHeadWeights{{1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f}};

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
// head values [cells][kHeadChannels] -> SH-0 slab (14 x gw*gh floats).
ds_status head_fwd(const float *h, const HeadGeom &g, float *slab, hipStream_t s);
// slab gradients -> head-value gradients (fp32 copy optional, bf16 always).
ds_status head_bwd(const float *h, const HeadGeom &g, const float *grad_slab, float *dh32,
                   __bf16 *dh, hipStream_t s);
// The head values whose splats best match a fitted slab (image position,
// depth, projected footprint, colour, opacity; each value through the
// inverse of its head function, clamped inside the function's range): the
// regression targets of supervised training, in the layout of h.
ds_status head_targets(const float *slab, const HeadGeom &g, float *target, hipStream_t s);
// dh = 2 w (h - target) / n over the head values (bf16), and
// loss += sum w (h - target)^2 / n (n = splats, w per head value).
ds_status head_regression(const float *h, const float *target, const HeadGeom &g,
                          const HeadWeights &w, __bf16 *dh, float *loss, hipStream_t s);

}  // namespace ds
