// Activation (log-scale/quaternion -> 3D covariance, SH -> colour, logit ->
// opacity) and world->NDC perspective projection, forward and backward.
//
// One thread per splat, every field a separate coalesced float array (SoA),
// SH degree a template parameter so the band loops unroll. The fused kernels
// keep the Splat3D intermediate in registers instead of HBM.
//
// Reference semantics:
//   covForwardKernel            src/layers/gauss_activ_layer.cu:37-170
//   covBackwardKernel           src/layers/gauss_activ_layer.cu:198-407
//   perspProjectForwardKernel   src/layers/persp_project_layer.cu:40-149
//   perspProjectBackwardKernel  src/layers/persp_project_layer.cu:173-267
#include <float.h>

#include "../ds_device.hpp"
#include "../ds_internal.hpp"
#include "forward_start.hpp"

namespace ds {
namespace {

constexpr int kBlock = 256;

struct Cov3 {
    float xx, xy, xz, yy, yz, zz;
};

// Rotation, scale and M = R * diag(s) of one splat.
struct RS {
    Quat q;
    Mat3 R;
    float s[3];
    float M[3][3];
};

__device__ __forceinline__ RS make_rs(float lsx, float lsy, float lsz, float rw, float rx,
                                      float ry, float rz) {
    RS o;
    o.q = quat_normalize(rw, rx, ry, rz);
    o.R = quat_to_rot(o.q);
    o.s[0] = expf(lsx);
    o.s[1] = expf(lsy);
    o.s[2] = expf(lsz);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) o.M[r][c] = o.R.r[r][c] * o.s[c];
    return o;
}

__device__ __forceinline__ Cov3 cov_of(const RS &o) {
    const auto &M = o.M;
    Cov3 c;
    c.xx = M[0][0] * M[0][0] + M[0][1] * M[0][1] + M[0][2] * M[0][2];
    c.xy = M[0][0] * M[1][0] + M[0][1] * M[1][1] + M[0][2] * M[1][2];
    c.xz = M[0][0] * M[2][0] + M[0][1] * M[2][1] + M[0][2] * M[2][2];
    c.yy = M[1][0] * M[1][0] + M[1][1] * M[1][1] + M[1][2] * M[1][2];
    c.yz = M[1][0] * M[2][0] + M[1][1] * M[2][1] + M[1][2] * M[2][2];
    c.zz = M[2][0] * M[2][0] + M[2][1] * M[2][1] + M[2][2] * M[2][2];
    return c;
}

// Unclamped colour (SH sum + 0.5) of splat i.
template <int DEG>
__device__ __forceinline__ float3 sh_color_raw(const ds_gauss3d &g, int i, float px, float py,
                                               float pz, Cam3 cam) {
    float r = g.sh_dc_r[i] * SH_C0;
    float gg = g.sh_dc_g[i] * SH_C0;
    float b = g.sh_dc_b[i] * SH_C0;
    if constexpr (DEG > 0) {
        constexpr int NB = sh_bands(DEG);
        const int n = g.count;
        float3 v = view_dir(cam.x, cam.y, cam.z, px, py, pz);
        float Y[15];
        sh_basis_rest<DEG>(v, Y);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            r += Y[k] * g.sh_rest_r[k * n + i];
            gg += Y[k] * g.sh_rest_g[k * n + i];
            b += Y[k] * g.sh_rest_b[k * n + i];
        }
    }
    return make_float3(r + 0.5f, gg + 0.5f, b + 0.5f);
}

__device__ __forceinline__ float clamp01(float c) { return fminf(fmaxf(c, 0.f), 1.f); }
__device__ __forceinline__ float sigmoid(float l) { return 1.f / (1.f + expf(-l)); }

// dL/d(raw params) of the rotation/scale chain from the full-matrix dL/dCov3D.
struct RSGrad {
    float ds[3];
    float dq[4];  // w, x, y, z (raw quaternion)
};

__device__ __forceinline__ RSGrad rs_backward(const RS &o, const Cov3 &G) {
    const float Gm[3][3] = {{G.xx, G.xy, G.xz}, {G.xy, G.yy, G.yz}, {G.xz, G.yz, G.zz}};
    float dM[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            dM[r][c] = 2.f * (Gm[r][0] * o.M[0][c] + Gm[r][1] * o.M[1][c] + Gm[r][2] * o.M[2][c]);
    RSGrad out;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out.ds[c] = (dM[0][c] * o.R.r[0][c] + dM[1][c] * o.R.r[1][c] + dM[2][c] * o.R.r[2][c]) *
                    o.s[c];
    float dR[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dR[r][c] = dM[r][c] * o.s[c];
    const float w = o.q.w, x = o.q.x, y = o.q.y, z = o.q.z;
    // Partial derivatives of quat_to_rot with respect to (w, x, y, z).
    float dw = 2.f * (-z * dR[0][1] + y * dR[0][2] + z * dR[1][0] - x * dR[1][2] - y * dR[2][0] +
                      x * dR[2][1]);
    float dx = 2.f * (y * dR[0][1] + z * dR[0][2] + y * dR[1][0] - 2.f * x * dR[1][1] -
                      w * dR[1][2] + z * dR[2][0] + w * dR[2][1] - 2.f * x * dR[2][2]);
    float dy = 2.f * (-2.f * y * dR[0][0] + x * dR[0][1] + w * dR[0][2] + x * dR[1][0] +
                      z * dR[1][2] - w * dR[2][0] + z * dR[2][1] - 2.f * y * dR[2][2]);
    float dz = 2.f * (-2.f * z * dR[0][0] - w * dR[0][1] + x * dR[0][2] + w * dR[1][0] -
                      2.f * z * dR[1][1] + y * dR[1][2] + x * dR[2][0] + y * dR[2][1]);
    // Through q_hat = q / |q|: inv_norm * (I - q_hat q_hat^T) dq_hat.
    float dot = dw * w + dx * x + dy * y + dz * z;
    out.dq[0] = o.q.inv_norm * (dw - dot * w);
    out.dq[1] = o.q.inv_norm * (dx - dot * x);
    out.dq[2] = o.q.inv_norm * (dy - dot * y);
    out.dq[3] = o.q.inv_norm * (dz - dot * z);
    return out;
}

// Writes dL/dSH for splat i from the gated colour gradient.
template <int DEG>
__device__ __forceinline__ void sh_backward(const ds_gauss3d &gi, int i, float px, float py,
                                            float pz, Cam3 cam, float dR, float dG, float dB) {
    gi.sh_dc_r[i] = dR * SH_C0;
    gi.sh_dc_g[i] = dG * SH_C0;
    gi.sh_dc_b[i] = dB * SH_C0;
    if constexpr (DEG > 0) {
        constexpr int NB = sh_bands(DEG);
        const int n = gi.count;
        float3 v = view_dir(cam.x, cam.y, cam.z, px, py, pz);
        float Y[15];
        sh_basis_rest<DEG>(v, Y);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            gi.sh_rest_r[k * n + i] = dR * Y[k];
            gi.sh_rest_g[k * n + i] = dG * Y[k];
            gi.sh_rest_b[k * n + i] = dB * Y[k];
        }
    }
}

// Projection of one point: NDC position and the 2x3 Jacobian d(ndc_xy)/d(world).
struct Proj {
    float cx, cy, cz, cw;
    float ndc[3];
    float J[2][3];
};

__device__ __forceinline__ bool project_point(const PV &pv, float x, float y, float z,
                                              Proj &p) {
    const float *m = pv.m;
    p.cx = m[0] * x + m[4] * y + m[8] * z + m[12];
    p.cy = m[1] * x + m[5] * y + m[9] * z + m[13];
    p.cz = m[2] * x + m[6] * y + m[10] * z + m[14];
    p.cw = m[3] * x + m[7] * y + m[11] * z + m[15];
    if (p.cw <= 0.f) return false;
    const float iw = 1.f / p.cw;
    const float iw2 = iw * iw;
    p.ndc[0] = p.cx * iw;
    p.ndc[1] = p.cy * iw;
    p.ndc[2] = p.cz * iw;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.J[0][k] = (m[4 * k + 0] * p.cw - p.cx * m[3 + 4 * k]) * iw2;
        p.J[1][k] = (m[4 * k + 1] * p.cw - p.cy * m[3 + 4 * k]) * iw2;
    }
    return true;
}

// Sigma2D = J Sigma3D J^T (upper triangle).
__device__ __forceinline__ void cov2d_of(const Proj &p, const Cov3 &c, float &oxx, float &oxy,
                                         float &oyy) {
    const float S[3][3] = {{c.xx, c.xy, c.xz}, {c.xy, c.yy, c.yz}, {c.xz, c.yz, c.zz}};
    float JS[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            JS[r][k] = p.J[r][0] * S[0][k] + p.J[r][1] * S[1][k] + p.J[r][2] * S[2][k];
    oxx = p.J[0][0] * JS[0][0] + p.J[0][1] * JS[0][1] + p.J[0][2] * JS[0][2];
    oxy = p.J[0][0] * JS[1][0] + p.J[0][1] * JS[1][1] + p.J[0][2] * JS[1][2];
    oyy = p.J[1][0] * JS[1][0] + p.J[1][1] * JS[1][1] + p.J[1][2] * JS[1][2];
}

// dL/dSigma3D (full-matrix convention) = J^T H J with H = [[a, b/2], [b/2, c]].
__device__ __forceinline__ Cov3 cov3_grad(const Proj &p, float a, float b, float c) {
    const float h = 0.5f * b;
    const auto &J = p.J;
    Cov3 g;
    g.xx = J[0][0] * J[0][0] * a + J[0][0] * J[1][0] * b + J[1][0] * J[1][0] * c;
    g.xy = J[0][0] * J[0][1] * a + h * (J[0][0] * J[1][1] + J[1][0] * J[0][1]) +
           J[1][0] * J[1][1] * c;
    g.xz = J[0][0] * J[0][2] * a + h * (J[0][0] * J[1][2] + J[1][0] * J[0][2]) +
           J[1][0] * J[1][2] * c;
    g.yy = J[0][1] * J[0][1] * a + J[0][1] * J[1][1] * b + J[1][1] * J[1][1] * c;
    g.yz = J[0][1] * J[0][2] * a + h * (J[0][1] * J[1][2] + J[1][1] * J[0][2]) +
           J[1][1] * J[1][2] * c;
    g.zz = J[0][2] * J[0][2] * a + J[0][2] * J[1][2] * b + J[1][2] * J[1][2] * c;
    return g;
}

// d Y_k / d(x, y, z) of the non-DC basis polynomials (x, y, z independent).
template <int DEG>
__device__ __forceinline__ void sh_basis_rest_grad(float3 v, float dY[15][3]) {
    const float x = v.x, y = v.y, z = v.z;
    auto set = [&](int k, float a, float b, float c) {
        dY[k][0] = a;
        dY[k][1] = b;
        dY[k][2] = c;
    };
    if constexpr (DEG >= 1) {
        set(0, 0.f, -SH_C1, 0.f);
        set(1, 0.f, 0.f, SH_C1);
        set(2, -SH_C1, 0.f, 0.f);
    }
    if constexpr (DEG >= 2) {
        set(3, SH_C2_0 * y, SH_C2_0 * x, 0.f);
        set(4, 0.f, SH_C2_1 * z, SH_C2_1 * y);
        set(5, -2.f * SH_C2_2 * x, -2.f * SH_C2_2 * y, 4.f * SH_C2_2 * z);
        set(6, SH_C2_3 * z, 0.f, SH_C2_3 * x);
        set(7, 2.f * SH_C2_4 * x, -2.f * SH_C2_4 * y, 0.f);
    }
    if constexpr (DEG >= 3) {
        const float xx = x * x, yy = y * y, zz = z * z;
        set(8, 6.f * SH_C3_0 * x * y, SH_C3_0 * (3.f * xx - 3.f * yy), 0.f);
        set(9, SH_C3_1 * y * z, SH_C3_1 * x * z, SH_C3_1 * x * y);
        set(10, -2.f * SH_C3_2 * x * y, SH_C3_2 * (4.f * zz - xx - 3.f * yy),
            8.f * SH_C3_2 * y * z);
        set(11, -6.f * SH_C3_3 * x * z, -6.f * SH_C3_3 * y * z,
            SH_C3_3 * (6.f * zz - 3.f * xx - 3.f * yy));
        set(12, SH_C3_4 * (4.f * zz - 3.f * xx - yy), -2.f * SH_C3_4 * x * y,
            8.f * SH_C3_4 * x * z);
        set(13, 2.f * SH_C3_5 * x * z, -2.f * SH_C3_5 * y * z, SH_C3_5 * (xx - yy));
        set(14, SH_C3_6 * (3.f * xx - 3.f * yy), -6.f * SH_C3_6 * x * y, 0.f);
    }
}

// Position gradient through the SH view direction v = (cam - p)/|cam - p|,
// added to g[3] (exact mode; the reference treats v as constant).
template <int DEG>
__device__ __forceinline__ void view_dir_pos_grad(const ds_gauss3d &in, int i, float px,
                                                  float py, float pz, Cam3 cam, float dR,
                                                  float dG, float dB, float g[3]) {
    if constexpr (DEG > 0) {
        constexpr int NB = sh_bands(DEG);
        const int n = in.count;
        const float dx = cam.x - px, dy = cam.y - py, dz = cam.z - pz;
        const float inv = rsqrtf(fmaxf(dx * dx + dy * dy + dz * dz, 1e-12f));
        const float3 v = make_float3(dx * inv, dy * inv, dz * inv);
        float dY[15][3];
        sh_basis_rest_grad<DEG>(v, dY);
        float gv[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const float w = dR * in.sh_rest_r[k * n + i] + dG * in.sh_rest_g[k * n + i] +
                            dB * in.sh_rest_b[k * n + i];
            gv[0] += w * dY[k][0];
            gv[1] += w * dY[k][1];
            gv[2] += w * dY[k][2];
        }
        const float vg = v.x * gv[0] + v.y * gv[1] + v.z * gv[2];
        g[0] -= (gv[0] - vg * v.x) * inv;
        g[1] -= (gv[1] - vg * v.y) * inv;
        g[2] -= (gv[2] - vg * v.z) * inv;
    }
}

// Position gradient through J in Sigma2D = J Sigma3D J^T (exact mode; the
// reference propagates position only through the projected mean).
__device__ __forceinline__ void jacobian_pos_grad(const PV &pv, const Proj &p, const Cov3 &c,
                                                  float a, float b, float cc, float g[3]) {
    const float S[3][3] = {{c.xx, c.xy, c.xz}, {c.xy, c.yy, c.yz}, {c.xz, c.yz, c.zz}};
    const float G2[2][2] = {{a, 0.5f * b}, {0.5f * b, cc}};
    float JS[2][3], dJ[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            JS[r][k] = p.J[r][0] * S[0][k] + p.J[r][1] * S[1][k] + p.J[r][2] * S[2][k];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) dJ[r][k] = 2.f * (G2[r][0] * JS[0][k] + G2[r][1] * JS[1][k]);
    const float *m = pv.m;
    const float iw = 1.f / p.cw;
    const float iw2 = iw * iw;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float acc = 0.f;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int col = 0; col < 3; ++col) {
                const float dJdp =
                    (m[4 * col + r] * m[4 * k + 3] - m[4 * k + r] * m[4 * col + 3]) * iw2 -
                    2.f * p.J[r][col] * m[4 * k + 3] * iw;
                acc += dJ[r][col] * dJdp;
            }
        g[k] += acc;
    }
}

// ============================================================ kernels

template <int DEG>
__global__ __launch_bounds__(kBlock) void k_activ_fwd(ds_gauss3d in, ds_splat3d out, Cam3 cam) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    const float px = in.pos_x[i], py = in.pos_y[i], pz = in.pos_z[i];
    out.pos_x[i] = px;
    out.pos_y[i] = py;
    out.pos_z[i] = pz;
    RS o = make_rs(in.scale_x[i], in.scale_y[i], in.scale_z[i], in.rot_w[i], in.rot_x[i],
                   in.rot_y[i], in.rot_z[i]);
    Cov3 c = cov_of(o);
    out.cov_xx[i] = c.xx;
    out.cov_xy[i] = c.xy;
    out.cov_xz[i] = c.xz;
    out.cov_yy[i] = c.yy;
    out.cov_yz[i] = c.yz;
    out.cov_zz[i] = c.zz;
    float3 col = sh_color_raw<DEG>(in, i, px, py, pz, cam);
    out.color_r[i] = clamp01(col.x);
    out.color_g[i] = clamp01(col.y);
    out.color_b[i] = clamp01(col.z);
    out.opacity[i] = sigmoid(in.logit_opacity[i]);
}

template <int DEG, int MODE>
__global__ __launch_bounds__(kBlock) void k_activ_bwd(ds_gauss3d in, ds_splat3d out,
                                                      ds_splat3d go, ds_gauss3d gi, Cam3 cam) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    RS o = make_rs(in.scale_x[i], in.scale_y[i], in.scale_z[i], in.rot_w[i], in.rot_x[i],
                   in.rot_y[i], in.rot_z[i]);
    Cov3 G{go.cov_xx[i], go.cov_xy[i], go.cov_xz[i], go.cov_yy[i], go.cov_yz[i], go.cov_zz[i]};
    RSGrad g = rs_backward(o, G);
    gi.scale_x[i] = g.ds[0];
    gi.scale_y[i] = g.ds[1];
    gi.scale_z[i] = g.ds[2];
    gi.rot_w[i] = g.dq[0];
    gi.rot_x[i] = g.dq[1];
    gi.rot_y[i] = g.dq[2];
    gi.rot_z[i] = g.dq[3];
    const float cr = out.color_r[i], cg = out.color_g[i], cb = out.color_b[i];
    const float dR = (cr > 0.f && cr < 1.f) ? go.color_r[i] : 0.f;
    const float dG = (cg > 0.f && cg < 1.f) ? go.color_g[i] : 0.f;
    const float dB = (cb > 0.f && cb < 1.f) ? go.color_b[i] : 0.f;
    float px = 0.f, py = 0.f, pz = 0.f;
    if constexpr (DEG > 0) {
        px = in.pos_x[i];
        py = in.pos_y[i];
        pz = in.pos_z[i];
    }
    sh_backward<DEG>(gi, i, px, py, pz, cam, dR, dG, dB);
    float gp[3] = {go.pos_x[i], go.pos_y[i], go.pos_z[i]};
    if constexpr (MODE == DS_GRAD_EXACT) view_dir_pos_grad<DEG>(in, i, px, py, pz, cam, dR, dG, dB, gp);
    gi.pos_x[i] = gp[0];
    gi.pos_y[i] = gp[1];
    gi.pos_z[i] = gp[2];
    const float s = out.opacity[i];
    gi.logit_opacity[i] = go.opacity[i] * s * (1.f - s);
}

__global__ __launch_bounds__(kBlock) void k_project_fwd(ds_splat3d in, ds_splat2d out, PV pv) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    Proj p;
    if (!project_point(pv, in.pos_x[i], in.pos_y[i], in.pos_z[i], p)) {
        out.pos_z[i] = FLT_MAX;  // culled: behind the camera
        return;
    }
    out.pos_x[i] = p.ndc[0];
    out.pos_y[i] = p.ndc[1];
    out.pos_z[i] = p.ndc[2];
    Cov3 c{in.cov_xx[i], in.cov_xy[i], in.cov_xz[i], in.cov_yy[i], in.cov_yz[i], in.cov_zz[i]};
    float oxx, oxy, oyy;
    cov2d_of(p, c, oxx, oxy, oyy);
    out.cov_xx[i] = oxx;
    out.cov_xy[i] = oxy;
    out.cov_yy[i] = oyy;
    out.color_r[i] = in.color_r[i];
    out.color_g[i] = in.color_g[i];
    out.color_b[i] = in.color_b[i];
    out.opacity[i] = in.opacity[i];
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_project_bwd(ds_splat3d in, ds_splat2d go,
                                                        ds_splat3d gi, PV pv) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    Proj p;
    if (!project_point(pv, in.pos_x[i], in.pos_y[i], in.pos_z[i], p)) {
        gi.pos_x[i] = 0.f; gi.pos_y[i] = 0.f; gi.pos_z[i] = 0.f;
        gi.cov_xx[i] = 0.f; gi.cov_xy[i] = 0.f; gi.cov_xz[i] = 0.f;
        gi.cov_yy[i] = 0.f; gi.cov_yz[i] = 0.f; gi.cov_zz[i] = 0.f;
        gi.color_r[i] = 0.f; gi.color_g[i] = 0.f; gi.color_b[i] = 0.f;
        gi.opacity[i] = 0.f;
        return;
    }
    Cov3 g = cov3_grad(p, go.cov_xx[i], go.cov_xy[i], go.cov_yy[i]);
    gi.cov_xx[i] = g.xx;
    gi.cov_xy[i] = g.xy;
    gi.cov_xz[i] = g.xz;
    gi.cov_yy[i] = g.yy;
    gi.cov_yz[i] = g.yz;
    gi.cov_zz[i] = g.zz;
    const float gx = go.pos_x[i], gy = go.pos_y[i];
    float gp[3] = {gx * p.J[0][0] + gy * p.J[1][0], gx * p.J[0][1] + gy * p.J[1][1],
                   gx * p.J[0][2] + gy * p.J[1][2]};
    if constexpr (MODE == DS_GRAD_EXACT) {
        Cov3 c{in.cov_xx[i], in.cov_xy[i], in.cov_xz[i], in.cov_yy[i], in.cov_yz[i], in.cov_zz[i]};
        jacobian_pos_grad(pv, p, c, go.cov_xx[i], go.cov_xy[i], go.cov_yy[i], gp);
    }
    gi.pos_x[i] = gp[0];
    gi.pos_y[i] = gp[1];
    gi.pos_z[i] = gp[2];
    gi.color_r[i] = go.color_r[i];
    gi.color_g[i] = go.color_g[i];
    gi.color_b[i] = go.color_b[i];
    gi.opacity[i] = go.opacity[i];
}

// Camera of splat i for the fused kernels: one camera for all splats, or a
// per-instance table indexed by each splat's instance (many fits at once).
struct CamUniform {
    Cam3 cam;
    PV pv;
    __device__ __forceinline__ void load(int, Cam3 &c, PV &p) const {
        c = cam;
        p = pv;
    }
};
struct CamTable {
    const float4 *tab;      // per instance: PV columns 0-3, then camera position
    const uint16_t *inst;   // per splat
    __device__ __forceinline__ void load(int i, Cam3 &c, PV &p) const {
        const float4 *r = tab + kCamRecF4 * inst[i];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 v = r[q];
            p.m[4 * q + 0] = v.x;
            p.m[4 * q + 1] = v.y;
            p.m[4 * q + 2] = v.z;
            p.m[4 * q + 3] = v.w;
        }
        const float4 e = r[4];
        c = Cam3{e.x, e.y, e.z};
    }
};

template <int DEG, class CS, bool START = false>
__global__ __launch_bounds__(kBlock) void k_activ_project_fwd(ds_gauss3d in, ds_splat2d out,
                                                              CS cs, ForwardStart fs) {
    // a rasterizer forward's first-kernel work, shared by this grid's threads
    if constexpr (START) forward_start(fs, gtid(), gstride());
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    Cam3 cam;
    PV pv;
    cs.load(i, cam, pv);
    const float px = in.pos_x[i], py = in.pos_y[i], pz = in.pos_z[i];
    Proj p;
    if (!project_point(pv, px, py, pz, p)) {
        out.pos_z[i] = FLT_MAX;
        return;
    }
    RS o = make_rs(in.scale_x[i], in.scale_y[i], in.scale_z[i], in.rot_w[i], in.rot_x[i],
                   in.rot_y[i], in.rot_z[i]);
    Cov3 c = cov_of(o);
    float oxx, oxy, oyy;
    cov2d_of(p, c, oxx, oxy, oyy);
    out.pos_x[i] = p.ndc[0];
    out.pos_y[i] = p.ndc[1];
    out.pos_z[i] = p.ndc[2];
    out.cov_xx[i] = oxx;
    out.cov_xy[i] = oxy;
    out.cov_yy[i] = oyy;
    float3 col = sh_color_raw<DEG>(in, i, px, py, pz, cam);
    out.color_r[i] = clamp01(col.x);
    out.color_g[i] = clamp01(col.y);
    out.color_b[i] = clamp01(col.z);
    out.opacity[i] = sigmoid(in.logit_opacity[i]);
}

template <int DEG, int MODE, class CS, bool ADAM = false>
__global__ __launch_bounds__(kBlock) void k_activ_project_bwd(ds_gauss3d in, ds_splat2d go,
                                                              ds_gauss3d gi, CS cs,
                                                              AdamFields af) {
    // With ADAM: the step's learning rate per field and bias corrections
    __shared__ float s_lr[ADAM ? kAdamFields : 1];
    __shared__ float s_bc[2];
    if constexpr (ADAM) {
        const int32_t t = *af.step;
        if (threadIdx.x < kAdamFields)
            s_lr[threadIdx.x] = adam_lr(af.lr[threadIdx.x], af.decay[threadIdx.x], t);
        if (threadIdx.x == 0) {
            s_bc[0] = 1.f / (1.f - powf(af.beta1, (float)t));
            s_bc[1] = 1.f / (1.f - powf(af.beta2, (float)t));
        }
        __syncthreads();
    }
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= in.count) return;
    // The 14 per-splat fields' parameters and moments are loaded up front so
    // their latency overlaps the gradient computation (nothing before the
    // Adam step writes them).
    float ap[14], am[14], av[14];
    if constexpr (ADAM) {
#pragma unroll
        for (int f = 0; f < 14; ++f) {
            ap[f] = af.p[f][i];
            am[f] = af.m[f][i];
            av[f] = af.v[f][i];
        }
    }
    // Adam step of field f at element e of its slab, gradient g
    auto adam = [&](int f, int64_t e, float g) {
        float pp = af.p[f][e], mm = af.m[f][e], vv = af.v[f][e];
        adam_elem(pp, mm, vv, g, s_lr[f], s_bc[0], s_bc[1], af.beta1, af.beta2, af.eps);
        af.p[f][e] = pp;
        af.m[f][e] = mm;
        af.v[f][e] = vv;
    };
    // the same for one of the 14 prefetched fields (element i)
    auto adam14 = [&](int f, float g) {
        adam_elem(ap[f], am[f], av[f], g, s_lr[f], s_bc[0], s_bc[1], af.beta1, af.beta2, af.eps);
        af.p[f][i] = ap[f];
        af.m[f][i] = am[f];
        af.v[f][i] = av[f];
    };
    Cam3 cam;
    PV pv;
    cs.load(i, cam, pv);
    const float px = in.pos_x[i], py = in.pos_y[i], pz = in.pos_z[i];
    Proj p;
    if (!project_point(pv, px, py, pz, p)) {
        gi.pos_x[i] = 0.f; gi.pos_y[i] = 0.f; gi.pos_z[i] = 0.f;
        gi.scale_x[i] = 0.f; gi.scale_y[i] = 0.f; gi.scale_z[i] = 0.f;
        gi.rot_w[i] = 0.f; gi.rot_x[i] = 0.f; gi.rot_y[i] = 0.f; gi.rot_z[i] = 0.f;
        gi.logit_opacity[i] = 0.f;
        sh_backward<DEG>(gi, i, px, py, pz, cam, 0.f, 0.f, 0.f);
        if constexpr (ADAM) {  // a zero gradient still moves the moments
#pragma unroll
            for (int f = 0; f < 14; ++f) adam14(f, 0.f);
            constexpr int nb = sh_bands(DEG);
            for (int k = 0; k < nb; ++k)
                for (int c = 0; c < 3; ++c) adam(14 + c, (int64_t)k * in.count + i, 0.f);
        }
        return;
    }
    // projection backward
    const float ga = go.cov_xx[i], gb = go.cov_xy[i], gc = go.cov_yy[i];
    Cov3 G = cov3_grad(p, ga, gb, gc);
    const float gx = go.pos_x[i], gy = go.pos_y[i];
    float gp[3] = {gx * p.J[0][0] + gy * p.J[1][0], gx * p.J[0][1] + gy * p.J[1][1],
                   gx * p.J[0][2] + gy * p.J[1][2]};
    // activation backward
    RS o = make_rs(in.scale_x[i], in.scale_y[i], in.scale_z[i], in.rot_w[i], in.rot_x[i],
                   in.rot_y[i], in.rot_z[i]);
    if constexpr (MODE == DS_GRAD_EXACT) jacobian_pos_grad(pv, p, cov_of(o), ga, gb, gc, gp);
    RSGrad g = rs_backward(o, G);
    gi.scale_x[i] = g.ds[0];
    gi.scale_y[i] = g.ds[1];
    gi.scale_z[i] = g.ds[2];
    gi.rot_w[i] = g.dq[0];
    gi.rot_x[i] = g.dq[1];
    gi.rot_y[i] = g.dq[2];
    gi.rot_z[i] = g.dq[3];
    float3 col = sh_color_raw<DEG>(in, i, px, py, pz, cam);
    const float cr = clamp01(col.x), cg = clamp01(col.y), cb = clamp01(col.z);
    const float dR = (cr > 0.f && cr < 1.f) ? go.color_r[i] : 0.f;
    const float dG = (cg > 0.f && cg < 1.f) ? go.color_g[i] : 0.f;
    const float dB = (cb > 0.f && cb < 1.f) ? go.color_b[i] : 0.f;
    sh_backward<DEG>(gi, i, px, py, pz, cam, dR, dG, dB);
    if constexpr (MODE == DS_GRAD_EXACT) view_dir_pos_grad<DEG>(in, i, px, py, pz, cam, dR, dG, dB, gp);
    gi.pos_x[i] = gp[0];
    gi.pos_y[i] = gp[1];
    gi.pos_z[i] = gp[2];
    const float s = sigmoid(in.logit_opacity[i]);
    const float g_op = go.opacity[i] * s * (1.f - s);
    gi.logit_opacity[i] = g_op;
    if constexpr (ADAM) {
        // every read of this splat's parameters is done: step them in place
        adam14(0, gp[0]);
        adam14(1, gp[1]);
        adam14(2, gp[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) adam14(3 + c, g.ds[c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) adam14(6 + c, g.dq[c]);
        adam14(10, dR * SH_C0);
        adam14(11, dG * SH_C0);
        adam14(12, dB * SH_C0);
        adam14(13, g_op);
        constexpr int nb = sh_bands(DEG);
        const float *rest[3] = {gi.sh_rest_r, gi.sh_rest_g, gi.sh_rest_b};
        for (int k = 0; k < nb; ++k)
            for (int c = 0; c < 3; ++c) {
                const int64_t e = (int64_t)k * in.count + i;
                adam(14 + c, e, rest[c][e]);
            }
    }
}

inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

#define DS_DISPATCH_DEG(deg, KERNEL, ...)                                                    \
    switch (deg) {                                                                          \
        case 0: hipLaunchKernelGGL(KERNEL<0>, __VA_ARGS__); break;                          \
        case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                          \
        case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                          \
        default: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                         \
    }

#define DS_DISPATCH_DEG_MODE(deg, mode, KERNEL, ...)                                         \
    if ((mode) == DS_GRAD_REFERENCE) {                                                      \
        switch (deg) {                                                                      \
            case 0: hipLaunchKernelGGL((KERNEL<0, DS_GRAD_REFERENCE>), __VA_ARGS__); break; \
            case 1: hipLaunchKernelGGL((KERNEL<1, DS_GRAD_REFERENCE>), __VA_ARGS__); break; \
            case 2: hipLaunchKernelGGL((KERNEL<2, DS_GRAD_REFERENCE>), __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<3, DS_GRAD_REFERENCE>), __VA_ARGS__); break;\
        }                                                                                   \
    } else {                                                                                \
        switch (deg) {                                                                      \
            case 0: hipLaunchKernelGGL((KERNEL<0, DS_GRAD_EXACT>), __VA_ARGS__); break;     \
            case 1: hipLaunchKernelGGL((KERNEL<1, DS_GRAD_EXACT>), __VA_ARGS__); break;     \
            case 2: hipLaunchKernelGGL((KERNEL<2, DS_GRAD_EXACT>), __VA_ARGS__); break;     \
            default: hipLaunchKernelGGL((KERNEL<3, DS_GRAD_EXACT>), __VA_ARGS__); break;    \
        }                                                                                   \
    }

ds_status launch_activ_fwd(const ds_gauss3d &in, const ds_splat3d &out, int deg, Cam3 cam,
                           hipStream_t s) {
    if (in.count <= 0) return DS_OK;
    DS_DISPATCH_DEG(deg, k_activ_fwd, dim3(grid_for(in.count)), dim3(kBlock), 0, s, in, out, cam);
    DS_CHECK_LAUNCH("activ_fwd");
    return DS_OK;
}

ds_status launch_activ_bwd(const ds_gauss3d &in, const ds_splat3d &out, const ds_splat3d &go,
                           const ds_gauss3d &gi, int deg, Cam3 cam, int mode, hipStream_t s) {
    if (in.count <= 0) return DS_OK;
    DS_DISPATCH_DEG_MODE(deg, mode, k_activ_bwd, dim3(grid_for(in.count)), dim3(kBlock), 0, s,
                         in, out, go, gi, cam);
    DS_CHECK_LAUNCH("activ_bwd");
    return DS_OK;
}

ds_status launch_project_fwd(const ds_splat3d &in, const ds_splat2d &out, const PV &pv,
                             hipStream_t s) {
    if (in.count <= 0) return DS_OK;
    hipLaunchKernelGGL(k_project_fwd, dim3(grid_for(in.count)), dim3(kBlock), 0, s, in, out, pv);
    DS_CHECK_LAUNCH("project_fwd");
    return DS_OK;
}

ds_status launch_project_bwd(const ds_splat3d &in, const ds_splat2d &go, const ds_splat3d &gi,
                             const PV &pv, int mode, hipStream_t s) {
    if (in.count <= 0) return DS_OK;
    if (mode == DS_GRAD_REFERENCE)
        hipLaunchKernelGGL(k_project_bwd<DS_GRAD_REFERENCE>, dim3(grid_for(in.count)),
                           dim3(kBlock), 0, s, in, go, gi, pv);
    else
        hipLaunchKernelGGL(k_project_bwd<DS_GRAD_EXACT>, dim3(grid_for(in.count)), dim3(kBlock),
                           0, s, in, go, gi, pv);
    DS_CHECK_LAUNCH("project_bwd");
    return DS_OK;
}

template <class CS, bool START = false>
static ds_status activ_project_fwd(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                   const CS &cs, hipStream_t s,
                                   const ForwardStart &fs = ForwardStart{}) {
    if (in.count <= 0) return DS_OK;
    const dim3 grid(grid_for(in.count)), block(kBlock);
    switch (deg) {
        case 0: hipLaunchKernelGGL((k_activ_project_fwd<0, CS, START>), grid, block, 0, s, in, out, cs, fs); break;
        case 1: hipLaunchKernelGGL((k_activ_project_fwd<1, CS, START>), grid, block, 0, s, in, out, cs, fs); break;
        case 2: hipLaunchKernelGGL((k_activ_project_fwd<2, CS, START>), grid, block, 0, s, in, out, cs, fs); break;
        default: hipLaunchKernelGGL((k_activ_project_fwd<3, CS, START>), grid, block, 0, s, in, out, cs, fs); break;
    }
    DS_CHECK_LAUNCH("activ_project_fwd");
    return DS_OK;
}

template <int MODE, class CS, bool ADAM>
static void activ_project_bwd_mode(const ds_gauss3d &in, const ds_splat2d &go,
                                   const ds_gauss3d &gi, int deg, const CS &cs,
                                   const AdamFields &af, hipStream_t s) {
    const dim3 grid(grid_for(in.count)), block(kBlock);
    switch (deg) {
        case 0: hipLaunchKernelGGL((k_activ_project_bwd<0, MODE, CS, ADAM>), grid, block, 0, s, in, go, gi, cs, af); break;
        case 1: hipLaunchKernelGGL((k_activ_project_bwd<1, MODE, CS, ADAM>), grid, block, 0, s, in, go, gi, cs, af); break;
        case 2: hipLaunchKernelGGL((k_activ_project_bwd<2, MODE, CS, ADAM>), grid, block, 0, s, in, go, gi, cs, af); break;
        default: hipLaunchKernelGGL((k_activ_project_bwd<3, MODE, CS, ADAM>), grid, block, 0, s, in, go, gi, cs, af); break;
    }
}

template <class CS, bool ADAM = false>
static ds_status activ_project_bwd(const ds_gauss3d &in, const ds_splat2d &go,
                                   const ds_gauss3d &gi, int deg, const CS &cs, int mode,
                                   hipStream_t s, const AdamFields &af = AdamFields{}) {
    if (in.count <= 0) return DS_OK;
    if (mode == DS_GRAD_REFERENCE)
        activ_project_bwd_mode<DS_GRAD_REFERENCE, CS, ADAM>(in, go, gi, deg, cs, af, s);
    else
        activ_project_bwd_mode<DS_GRAD_EXACT, CS, ADAM>(in, go, gi, deg, cs, af, s);
    DS_CHECK_LAUNCH("activ_project_bwd");
    return DS_OK;
}

ds_status launch_activ_project_fwd(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                   Cam3 cam, const PV &pv, hipStream_t s) {
    return activ_project_fwd(in, out, deg, CamUniform{cam, pv}, s);
}

ds_status launch_activ_project_bwd(const ds_gauss3d &in, const ds_splat2d &go,
                                   const ds_gauss3d &gi, int deg, Cam3 cam, const PV &pv,
                                   int mode, hipStream_t s) {
    return activ_project_bwd(in, go, gi, deg, CamUniform{cam, pv}, mode, s);
}

ds_status launch_activ_project_fwd_start(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                         Cam3 cam, const PV &pv, const ForwardStart &fs,
                                         hipStream_t s) {
    if (in.count <= 0) return set_error(DS_ERR_INVALID_ARG, "forward start needs splats");
    return activ_project_fwd<CamUniform, true>(in, out, deg, CamUniform{cam, pv}, s, fs);
}

ds_status launch_activ_project_fwd_start_multi(const ds_gauss3d &in, const ds_splat2d &out,
                                               int deg, const float4 *cam_tab,
                                               const uint16_t *splat_inst,
                                               const ForwardStart &fs, hipStream_t s) {
    if (in.count <= 0) return set_error(DS_ERR_INVALID_ARG, "forward start needs splats");
    return activ_project_fwd<CamTable, true>(in, out, deg, CamTable{cam_tab, splat_inst}, s, fs);
}

ds_status launch_activ_project_bwd_adam_multi(const ds_gauss3d &in, const ds_splat2d &go,
                                              const ds_gauss3d &gi, int deg,
                                              const float4 *cam_tab, const uint16_t *splat_inst,
                                              int mode, const AdamFields &af, hipStream_t s) {
    if (!af.step) return set_error(DS_ERR_INVALID_ARG, "fused Adam step without a step counter");
    return activ_project_bwd<CamTable, true>(in, go, gi, deg, CamTable{cam_tab, splat_inst},
                                             mode, s, af);
}

ds_status launch_activ_project_bwd_adam(const ds_gauss3d &in, const ds_splat2d &go,
                                        const ds_gauss3d &gi, int deg, Cam3 cam, const PV &pv,
                                        int mode, const AdamFields &af, hipStream_t s) {
    if (!af.step) return set_error(DS_ERR_INVALID_ARG, "fused Adam step without a step counter");
    return activ_project_bwd<CamUniform, true>(in, go, gi, deg, CamUniform{cam, pv}, mode, s,
                                               af);
}

ds_status launch_activ_project_fwd_multi(const ds_gauss3d &in, const ds_splat2d &out, int deg,
                                         const float4 *cam_tab, const uint16_t *splat_inst,
                                         hipStream_t s) {
    return activ_project_fwd(in, out, deg, CamTable{cam_tab, splat_inst}, s);
}

ds_status launch_activ_project_bwd_multi(const ds_gauss3d &in, const ds_splat2d &go,
                                         const ds_gauss3d &gi, int deg, const float4 *cam_tab,
                                         const uint16_t *splat_inst, int mode, hipStream_t s) {
    return activ_project_bwd(in, go, gi, deg, CamTable{cam_tab, splat_inst}, mode, s);
}

}  // namespace ds
