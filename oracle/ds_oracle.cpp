// CPU oracle for the Gaussian-splatting hot path -- TEST INFRASTRUCTURE ONLY.
//
// This file is the checker: tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the engine itself never does.
//
// It restates the reference's kernels as plain loops, templated on the
// scalar type:
//   T = float   "transcription": the reference's fp32 operation order with
//               no FP contraction (build with -ffp-contract=off). Integer
//               outputs (tile rectangles, keys, sorted order, ranges) are
//               bit-exact with the engine given the same fp32 inputs. Also the
//               CPU baseline (OpenMP over splats, tiles and parameters).
//   T = double  FP64 oracle for images, gradients and Adam updates.
//
// Compositing records, per pixel, the smallest relative margin of any
// discrete decision it took (alpha >= 1/255, o*g <= 0.99, T <= 1e-4), so a
// test can tell decision flips between fp32 and FP64 from real errors. The
// backward can bound, per gradient element, how much of it comes from such
// pixels (the terms a decision flip there may change).
//
// Reference functions restated (file:line in the reference tree):
//   act_fwd        src/layers/gauss_activ_layer.cu:37-170
//   act_bwd        src/layers/gauss_activ_layer.cu:198-407
//   quat/rot/view  src/cuda/math_ops.cuh:34-101
//   SH constants   src/utils/sh_consts.h:7-35
//   proj_fwd       src/layers/persp_project_layer.cu:40-149
//   proj_bwd       src/layers/persp_project_layer.cu:173-267
//   make_key       src/layers/tile_rasterize_layer.cu:30-38
//   tile_assign    src/layers/tile_rasterize_layer.cu:52-144
//   sort           cub::DeviceRadixSort::SortPairs, tile_rasterize_layer.cu:667-683
//                  (stable ascending by 64-bit key)
//   tile_ranges    src/layers/tile_rasterize_layer.cu:148-162
//   comp_fwd       src/layers/tile_rasterize_layer.cu:192-309
//   comp_bwd       src/layers/tile_rasterize_layer.cu:356-578
//   mse            src/layers/mse_loss_layer.cu:22-67
//   adam           src/optimizers/adam.cu:12-37, 59-73
//   step order     src/pipelines/image_fitter.cpp:103-112, pipeline.h:11-27
//   densify        no reference code (README.md:10 lists it as a TODO): Kerbl
//                  et al., SIGGRAPH 2023, sec. 5.2 (clone / split / prune)
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

template <typename T>
struct K {
    static constexpr T C0 = (T)0.28209479177387814;
    static constexpr T C1 = (T)0.4886025119029199;
    static constexpr T C2_0 = (T)1.0925484305920792;
    static constexpr T C2_1 = (T)-1.0925484305920792;
    static constexpr T C2_2 = (T)0.3153915652525200;
    static constexpr T C2_3 = (T)-1.0925484305920792;
    static constexpr T C2_4 = (T)0.5462742152960396;
    static constexpr T C3_0 = (T)-0.5900435899266435;
    static constexpr T C3_1 = (T)2.8906114426405540;
    static constexpr T C3_2 = (T)-0.4570457994644658;
    static constexpr T C3_3 = (T)0.3731763325901154;
    static constexpr T C3_4 = (T)-0.4570457994644658;
    static constexpr T C3_5 = (T)1.4453057213202770;
    static constexpr T C3_6 = (T)-0.5900435899266435;
};

inline int bands_of(int d) { return d <= 0 ? 0 : d == 1 ? 3 : d == 2 ? 8 : 15; }

template <typename T>
inline T rsqrt_(T x) {
    return (T)1 / std::sqrt(x);
}

// Field pointers of a parameter slab (same layout as the engine's).
template <typename T>
struct PSlab {
    T *f[13];
    T *rr, *rg, *rb, *logit;
    PSlab(T *p, int n, int nb) {
        for (int k = 0; k < 13; ++k) f[k] = p + (size_t)k * n;
        rr = p + 13 * (size_t)n;
        rg = rr + (size_t)nb * n;
        rb = rg + (size_t)nb * n;
        logit = rb + (size_t)nb * n;
    }
};

// SH basis for the 15 non-DC bands, math_ops/gauss_activ order.
template <typename T>
void sh_rest_basis(int deg, T vx, T vy, T vz, T *Y) {
    using C = K<T>;
    if (deg >= 1) {
        Y[0] = -C::C1 * vy;
        Y[1] = C::C1 * vz;
        Y[2] = -C::C1 * vx;
    }
    if (deg >= 2) {
        T xx = vx * vx, yy = vy * vy, zz = vz * vz, xy = vx * vy, yz = vy * vz, xz = vx * vz;
        Y[3] = C::C2_0 * xy;
        Y[4] = C::C2_1 * yz;
        Y[5] = C::C2_2 * (2 * zz - xx - yy);
        Y[6] = C::C2_3 * xz;
        Y[7] = C::C2_4 * (xx - yy);
        if (deg >= 3) {
            Y[8] = C::C3_0 * vy * (3 * xx - yy);
            Y[9] = C::C3_1 * xy * vz;
            Y[10] = C::C3_2 * vy * (4 * zz - xx - yy);
            Y[11] = C::C3_3 * vz * (2 * zz - 3 * xx - 3 * yy);
            Y[12] = C::C3_4 * vx * (4 * zz - xx - yy);
            Y[13] = C::C3_5 * vz * (xx - yy);
            Y[14] = C::C3_6 * vx * (xx - 3 * yy);
        }
    }
}

// d Y_k / d(x, y, z) of the 15 non-DC basis polynomials, with x, y, z
// treated as independent (the caller projects onto the unit sphere).
template <typename T>
void sh_rest_basis_grad(int deg, T x, T y, T z, T dY[15][3]) {
    using C = K<T>;
    auto set = [&](int k, T a, T b, T c) { dY[k][0] = a; dY[k][1] = b; dY[k][2] = c; };
    if (deg >= 1) {
        set(0, 0, -C::C1, 0);
        set(1, 0, 0, C::C1);
        set(2, -C::C1, 0, 0);
    }
    if (deg >= 2) {
        set(3, C::C2_0 * y, C::C2_0 * x, 0);
        set(4, 0, C::C2_1 * z, C::C2_1 * y);
        set(5, -2 * C::C2_2 * x, -2 * C::C2_2 * y, 4 * C::C2_2 * z);
        set(6, C::C2_3 * z, 0, C::C2_3 * x);
        set(7, 2 * C::C2_4 * x, -2 * C::C2_4 * y, 0);
    }
    if (deg >= 3) {
        const T xx = x * x, yy = y * y, zz = z * z;
        set(8, 6 * C::C3_0 * x * y, C::C3_0 * (3 * xx - 3 * yy), 0);
        set(9, C::C3_1 * y * z, C::C3_1 * x * z, C::C3_1 * x * y);
        set(10, -2 * C::C3_2 * x * y, C::C3_2 * (4 * zz - xx - 3 * yy), 8 * C::C3_2 * y * z);
        set(11, -6 * C::C3_3 * x * z, -6 * C::C3_3 * y * z, C::C3_3 * (6 * zz - 3 * xx - 3 * yy));
        set(12, C::C3_4 * (4 * zz - 3 * xx - yy), -2 * C::C3_4 * x * y, 8 * C::C3_4 * x * z);
        set(13, 2 * C::C3_5 * x * z, -2 * C::C3_5 * y * z, C::C3_5 * (xx - yy));
        set(14, C::C3_6 * (3 * xx - 3 * yy), -6 * C::C3_6 * x * y, 0);
    }
}

template <typename T>
struct RSq {
    T q[4], inv;  // unit quaternion w,x,y,z and 1/|q_raw|
    T R[3][3];
    T s[3];
    T M[3][3];
};

// normalizeQuat + quatToRotMat + M = R diag(exp(log-scale)).
template <typename T>
RSq<T> make_rs(T lsx, T lsy, T lsz, T w, T x, T y, T z) {
    RSq<T> o;
    T inv = rsqrt_<T>(std::max(w * w + x * x + y * y + z * z, (T)1e-12));
    o.inv = inv;
    o.q[0] = w * inv;
    o.q[1] = x * inv;
    o.q[2] = y * inv;
    o.q[3] = z * inv;
    T qw = o.q[0], qx = o.q[1], qy = o.q[2], qz = o.q[3];
    T R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)},
                 {2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)},
                 {2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)}};
    std::memcpy(o.R, R, sizeof(R));
    o.s[0] = std::exp(lsx);
    o.s[1] = std::exp(lsy);
    o.s[2] = std::exp(lsz);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.M[r][c] = o.R[r][c] * o.s[c];
    return o;
}

// ---------------------------------------------------------------- activation
template <typename T>
void act_fwd(const T *params, int n, int deg, const T cam[3], T *s3) {
    const int nb = bands_of(deg);
    PSlab<T> p(const_cast<T *>(params), n, nb);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        T px = p.f[0][i], py = p.f[1][i], pz = p.f[2][i];
        s3[0 * (size_t)n + i] = px;
        s3[1 * (size_t)n + i] = py;
        s3[2 * (size_t)n + i] = pz;
        RSq<T> o = make_rs<T>(p.f[3][i], p.f[4][i], p.f[5][i], p.f[6][i], p.f[7][i], p.f[8][i],
                              p.f[9][i]);
        auto &M = o.M;
        s3[3 * (size_t)n + i] = M[0][0] * M[0][0] + M[0][1] * M[0][1] + M[0][2] * M[0][2];
        s3[4 * (size_t)n + i] = M[0][0] * M[1][0] + M[0][1] * M[1][1] + M[0][2] * M[1][2];
        s3[5 * (size_t)n + i] = M[0][0] * M[2][0] + M[0][1] * M[2][1] + M[0][2] * M[2][2];
        s3[6 * (size_t)n + i] = M[1][0] * M[1][0] + M[1][1] * M[1][1] + M[1][2] * M[1][2];
        s3[7 * (size_t)n + i] = M[1][0] * M[2][0] + M[1][1] * M[2][1] + M[1][2] * M[2][2];
        s3[8 * (size_t)n + i] = M[2][0] * M[2][0] + M[2][1] * M[2][1] + M[2][2] * M[2][2];
        T col[3] = {p.f[10][i] * K<T>::C0, p.f[11][i] * K<T>::C0, p.f[12][i] * K<T>::C0};
        if (deg > 0) {
            T dx = cam[0] - px, dy = cam[1] - py, dz = cam[2] - pz;
            T inv = rsqrt_<T>(std::max(dx * dx + dy * dy + dz * dz, (T)1e-12));
            T Y[15];
            sh_rest_basis<T>(deg, dx * inv, dy * inv, dz * inv, Y);
            const T *rest[3] = {p.rr, p.rg, p.rb};
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < nb; ++k) col[c] += Y[k] * rest[c][(size_t)k * n + i];
        }
        for (int c = 0; c < 3; ++c)
            s3[(9 + c) * (size_t)n + i] = std::min(std::max(col[c] + (T)0.5, (T)0), (T)1);
        s3[12 * (size_t)n + i] = (T)1 / ((T)1 + std::exp(-p.logit[i]));
    }
}

template <typename T>
// grad_mode 0 adds the position gradient through the SH view direction,
// which the reference treats as constant (gauss_activ_layer.cu:201-203).
void act_bwd(const T *params, const T *s3, const T *gs3, int n, int deg, const T cam[3],
             T *gparams, int grad_mode) {
    const int nb = bands_of(deg);
    PSlab<T> p(const_cast<T *>(params), n, nb);
    PSlab<T> g(gparams, n, nb);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        auto G3 = [&](int f) { return gs3[(size_t)f * n + i]; };
        for (int k = 0; k < 3; ++k) g.f[k][i] = G3(k);
        RSq<T> o = make_rs<T>(p.f[3][i], p.f[4][i], p.f[5][i], p.f[6][i], p.f[7][i], p.f[8][i],
                              p.f[9][i]);
        const T Gm[3][3] = {{G3(3), G3(4), G3(5)}, {G3(4), G3(6), G3(7)}, {G3(5), G3(7), G3(8)}};
        T dM[3][3], dR[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                dM[r][c] = 2 * (Gm[r][0] * o.M[0][c] + Gm[r][1] * o.M[1][c] + Gm[r][2] * o.M[2][c]);
        for (int c = 0; c < 3; ++c)
            g.f[3 + c][i] =
                (dM[0][c] * o.R[0][c] + dM[1][c] * o.R[1][c] + dM[2][c] * o.R[2][c]) * o.s[c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) dR[r][c] = dM[r][c] * o.s[c];
        const T w = o.q[0], x = o.q[1], y = o.q[2], z = o.q[3];
        T dw = 2 * (dR[1][0] * z - dR[2][0] * y + dR[0][1] * (-z) + dR[2][1] * x + dR[0][2] * y +
                    dR[1][2] * (-x));
        T dx = 2 * (dR[1][0] * y + dR[2][0] * z + dR[0][1] * y + dR[1][1] * (-2 * x) +
                    dR[2][1] * w + dR[0][2] * z + dR[1][2] * (-w) + dR[2][2] * (-2 * x));
        T dy = 2 * (dR[0][0] * (-2 * y) + dR[1][0] * x + dR[2][0] * (-w) + dR[0][1] * x +
                    dR[1][2] * z + dR[2][1] * z + dR[2][2] * (-2 * y) + dR[0][2] * w);
        T dz = 2 * (dR[0][0] * (-2 * z) + dR[1][0] * w + dR[2][0] * x + dR[0][1] * (-w) +
                    dR[1][1] * (-2 * z) + dR[1][2] * y + dR[2][1] * y + dR[0][2] * x);
        T dot = dw * w + dx * x + dy * y + dz * z;
        g.f[6][i] = o.inv * (dw - dot * w);
        g.f[7][i] = o.inv * (dx - dot * x);
        g.f[8][i] = o.inv * (dy - dot * y);
        g.f[9][i] = o.inv * (dz - dot * z);
        T gc[3];
        for (int c = 0; c < 3; ++c) {
            T col = s3[(9 + c) * (size_t)n + i];
            gc[c] = (col > 0 && col < 1) ? G3(9 + c) : (T)0;
            g.f[10 + c][i] = gc[c] * K<T>::C0;
        }
        if (deg > 0) {
            T px = p.f[0][i], py = p.f[1][i], pz = p.f[2][i];
            T ddx = cam[0] - px, ddy = cam[1] - py, ddz = cam[2] - pz;
            T inv = rsqrt_<T>(std::max(ddx * ddx + ddy * ddy + ddz * ddz, (T)1e-12));
            T Y[15];
            sh_rest_basis<T>(deg, ddx * inv, ddy * inv, ddz * inv, Y);
            T *rest[3] = {g.rr, g.rg, g.rb};
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < nb; ++k) rest[c][(size_t)k * n + i] = gc[c] * Y[k];
            if (grad_mode == 0) {
                const T v[3] = {ddx * inv, ddy * inv, ddz * inv};
                T dY[15][3];
                sh_rest_basis_grad<T>(deg, v[0], v[1], v[2], dY);
                const T *coef[3] = {p.rr, p.rg, p.rb};
                T gv[3] = {0, 0, 0};
                for (int c = 0; c < 3; ++c)
                    for (int k = 0; k < nb; ++k)
                        for (int a = 0; a < 3; ++a)
                            gv[a] += gc[c] * coef[c][(size_t)k * n + i] * dY[k][a];
                const T vg = v[0] * gv[0] + v[1] * gv[1] + v[2] * gv[2];
                for (int a = 0; a < 3; ++a) g.f[a][i] -= (gv[a] - vg * v[a]) * inv;
            }
        }
        T s = s3[12 * (size_t)n + i];
        g.logit[i] = G3(12) * s * (1 - s);
    }
}

// ---------------------------------------------------------------- projection
template <typename T>
bool proj_point(const T *m, T x, T y, T z, T ndc[3], T J[2][3]) {
    T cx = m[0] * x + m[4] * y + m[8] * z + m[12];
    T cy = m[1] * x + m[5] * y + m[9] * z + m[13];
    T cz = m[2] * x + m[6] * y + m[10] * z + m[14];
    T cw = m[3] * x + m[7] * y + m[11] * z + m[15];
    if (cw <= 0) return false;
    T iw = 1 / cw, iw2 = iw * iw;
    ndc[0] = cx * iw;
    ndc[1] = cy * iw;
    ndc[2] = cz * iw;
    for (int k = 0; k < 3; ++k) {
        J[0][k] = (m[4 * k + 0] * cw - cx * m[4 * k + 3]) * iw2;
        J[1][k] = (m[4 * k + 1] * cw - cy * m[4 * k + 3]) * iw2;
    }
    return true;
}

template <typename T>
void proj_fwd(const T *s3, int n, const T pv[16], T *s2) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        auto S = [&](int f) { return s3[(size_t)f * n + i]; };
        T ndc[3], J[2][3];
        if (!proj_point<T>(pv, S(0), S(1), S(2), ndc, J)) {
            s2[2 * (size_t)n + i] = (T)FLT_MAX;
            continue;
        }
        for (int k = 0; k < 3; ++k) s2[(size_t)k * n + i] = ndc[k];
        const T C[3][3] = {{S(3), S(4), S(5)}, {S(4), S(6), S(7)}, {S(5), S(7), S(8)}};
        T JS[2][3];
        for (int r = 0; r < 2; ++r)
            for (int k = 0; k < 3; ++k)
                JS[r][k] = J[r][0] * C[0][k] + J[r][1] * C[1][k] + J[r][2] * C[2][k];
        s2[3 * (size_t)n + i] = J[0][0] * JS[0][0] + J[0][1] * JS[0][1] + J[0][2] * JS[0][2];
        s2[4 * (size_t)n + i] = J[0][0] * JS[1][0] + J[0][1] * JS[1][1] + J[0][2] * JS[1][2];
        s2[5 * (size_t)n + i] = J[1][0] * JS[1][0] + J[1][1] * JS[1][1] + J[1][2] * JS[1][2];
        for (int k = 0; k < 4; ++k) s2[(size_t)(6 + k) * n + i] = S(9 + k);
    }
}

// grad_mode 0 adds the position gradient through the Jacobian of the
// projected covariance, which the reference omits (persp_project_layer.cu:254-260).
template <typename T>
void proj_bwd(const T *s3, const T *gs2, int n, const T pv[16], T *gs3, int grad_mode) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        auto G = [&](int f) { return gs2[(size_t)f * n + i]; };
        T ndc[3], J[2][3];
        if (!proj_point<T>(pv, s3[i], s3[(size_t)n + i], s3[2 * (size_t)n + i], ndc, J)) {
            for (int f = 0; f < 13; ++f) gs3[(size_t)f * n + i] = 0;
            continue;
        }
        T a = G(3), b = G(4), c = G(5), h = b / 2;
        T *o = gs3;
        o[3 * (size_t)n + i] = J[0][0] * J[0][0] * a + J[0][0] * J[1][0] * b + J[1][0] * J[1][0] * c;
        o[4 * (size_t)n + i] = J[0][0] * J[0][1] * a + h * (J[0][0] * J[1][1] + J[1][0] * J[0][1]) +
                               J[1][0] * J[1][1] * c;
        o[5 * (size_t)n + i] = J[0][0] * J[0][2] * a + h * (J[0][0] * J[1][2] + J[1][0] * J[0][2]) +
                               J[1][0] * J[1][2] * c;
        o[6 * (size_t)n + i] = J[0][1] * J[0][1] * a + J[0][1] * J[1][1] * b + J[1][1] * J[1][1] * c;
        o[7 * (size_t)n + i] = J[0][1] * J[0][2] * a + h * (J[0][1] * J[1][2] + J[1][1] * J[0][2]) +
                               J[1][1] * J[1][2] * c;
        o[8 * (size_t)n + i] = J[0][2] * J[0][2] * a + J[0][2] * J[1][2] * b + J[1][2] * J[1][2] * c;
        T gx = G(0), gy = G(1);
        o[0 * (size_t)n + i] = gx * J[0][0] + gy * J[1][0];
        o[1 * (size_t)n + i] = gx * J[0][1] + gy * J[1][1];
        o[2 * (size_t)n + i] = gx * J[0][2] + gy * J[1][2];
        if (grad_mode == 0) {
            auto S = [&](int f) { return s3[(size_t)f * n + i]; };
            const T C3[3][3] = {{S(3), S(4), S(5)}, {S(4), S(6), S(7)}, {S(5), S(7), S(8)}};
            const T G2[2][2] = {{a, h}, {h, c}};
            T JS[2][3], dJ[2][3];
            for (int r = 0; r < 2; ++r)
                for (int k = 0; k < 3; ++k)
                    JS[r][k] = J[r][0] * C3[0][k] + J[r][1] * C3[1][k] + J[r][2] * C3[2][k];
            for (int r = 0; r < 2; ++r)
                for (int k = 0; k < 3; ++k) dJ[r][k] = 2 * (G2[r][0] * JS[0][k] + G2[r][1] * JS[1][k]);
            const T *m = pv;
            const T X = s3[i], Y = s3[(size_t)n + i], Z = s3[2 * (size_t)n + i];
            const T cw = m[3] * X + m[7] * Y + m[11] * Z + m[15];
            const T iw = 1 / cw, iw2 = iw * iw;
            for (int k = 0; k < 3; ++k) {
                T acc = 0;
                for (int r = 0; r < 2; ++r)
                    for (int col = 0; col < 3; ++col) {
                        const T dJdp = (m[4 * col + r] * m[4 * k + 3] - m[4 * k + r] * m[4 * col + 3]) * iw2 -
                                       2 * J[r][col] * m[4 * k + 3] * iw;
                        acc += dJ[r][col] * dJdp;
                    }
                o[(size_t)k * n + i] += acc;
            }
        }
        for (int k = 0; k < 4; ++k) o[(size_t)(9 + k) * n + i] = G(6 + k);
    }
}

// ---------------------------------------------------------------- tile assign (fp32)
inline uint32_t depth_bits(float z) {
    uint32_t u;
    std::memcpy(&u, &z, 4);
    return (u >> 31) ? ~u : (u ^ 0x80000000u);
}

inline int tile_floor(float v, int hi) {
    float f = std::floor(v);
    f = std::max(f, -1.0f);
    f = std::min(f, (float)hi);
    return (int)f;
}

struct Rect {
    int tx0, tx1, ty0, ty1;
    bool vis;
};

// tileAssignKernel's per-splat cull and rectangle, fp32, no contraction.
Rect tile_rect(const float *s2, int n, int i, int W, int H, int ntx, int nty) {
    Rect r{0, -1, 0, -1, false};
    const float x = s2[i], y = s2[(size_t)n + i], z = s2[2 * (size_t)n + i];
    if (!(std::fabs(x) <= 1.0f && std::fabs(y) <= 1.0f && std::fabs(z) <= 1.0f)) return r;
    const float cxx = s2[3 * (size_t)n + i], cxy = s2[4 * (size_t)n + i],
                cyy = s2[5 * (size_t)n + i];
    const float det = cxx * cyy - cxy * cxy;
    if (!(det > 0.f)) return r;
    const float ex = 3.f * std::sqrt(cxx);
    const float ey = 3.f * std::sqrt(cyy);
    const float pixel_ndc = std::max(2.f / (float)W, 2.f / (float)H);
    if (!(std::max(ex, ey) * 2.f >= pixel_ndc)) return r;
    const float fx = (float)W / 16.f, fy = (float)H / 16.f;
    r.tx0 = std::max(tile_floor((x - ex + 1.f) * 0.5f * fx, ntx), 0);
    r.tx1 = std::min(tile_floor((x + ex + 1.f) * 0.5f * fx, ntx), ntx - 1);
    r.ty0 = std::max(tile_floor((1.f - (y + ey)) * 0.5f * fy, nty), 0);
    r.ty1 = std::min(tile_floor((1.f - (y - ey)) * 0.5f * fy, nty), nty - 1);
    r.vis = (r.tx0 <= r.tx1 && r.ty0 <= r.ty1);
    return r;
}

// ---------------------------------------------------------------- compositing
constexpr double kMaxAlpha = 0.99;
constexpr double kAlphaThres = 1.0 / 255.0;
constexpr double kTThres = 0.0001;

template <typename T>
struct Rec {
    T x, y, a, b, c, o, idet, r, g, bl;
};

template <typename T>
Rec<T> rec_of(const T *s2, int n, uint32_t gi) {
    auto S = [&](int f) { return s2[(size_t)f * n + gi]; };
    Rec<T> R;
    T cxx = S(3), cxy = S(4), cyy = S(5);
    T det = cxx * cyy - cxy * cxy;
    T idet = 1 / det;
    R.x = S(0);
    R.y = S(1);
    R.a = cyy * idet;
    R.b = -cxy * idet;
    R.c = cxx * idet;
    R.o = S(9);
    R.idet = idet;
    R.r = S(6);
    R.g = S(7);
    R.bl = S(8);
    return R;
}

// Pixel centres in NDC are defined in fp32 (compositeFwdKernel,
// tile_rasterize_layer.cu:230-231), in every mode.
inline float pixel_x(int j, int W) { return 2.f * ((float)j + 0.5f) / (float)W - 1.f; }
inline float pixel_y(int i, int H) { return 1.f - 2.f * ((float)i + 0.5f) / (float)H; }

template <typename T>
inline double rel_margin(T v, double thr) {
    return std::fabs((double)v - thr) / thr;
}

template <typename T>
void comp_fwd(const T *s2, int n, const uint32_t *flat, const int32_t *ranges, int W, int H,
              T *rgb, T *Tout, int32_t *last, double *margin, int64_t *e_fwd) {
    const int ntx = (W + 15) / 16, nty = (H + 15) / 16;
    int64_t E = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : E)
    for (int tile = 0; tile < ntx * nty; ++tile) {
        const int ty = tile / ntx, tx = tile % ntx;
        const int rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        std::vector<Rec<T>> recs;
        recs.reserve(re > rs ? re - rs : 0);
        for (int k = rs; k < re; ++k) recs.push_back(rec_of<T>(s2, n, flat[k]));
        for (int ly = 0; ly < 16; ++ly) {
            for (int lx = 0; lx < 16; ++lx) {
                const int i = ty * 16 + ly, j = tx * 16 + lx;
                if (i >= H || j >= W) continue;
                const T px = pixel_x(j, W), py = pixel_y(i, H);
                T Tr = 1, cr = 0, cg = 0, cb = 0;
                int32_t cur = 0;
                double mg = 1e30;
                for (int k = rs; k < re; ++k) {
                    ++E;
                    const Rec<T> &R = recs[(size_t)(k - rs)];
                    const T dx = R.x - px, dy = R.y - py;
                    const T d2 = dx * dx * R.a + 2 * dx * dy * R.b + dy * dy * R.c;
                    const T sigma = (T)0.5 * d2;
                    if (sigma < 0) continue;
                    const T og = R.o * std::exp(-sigma);
                    mg = std::min(mg, rel_margin(og, kMaxAlpha));
                    const T alpha = std::min((T)kMaxAlpha, og);
                    mg = std::min(mg, rel_margin(alpha, kAlphaThres));
                    if (alpha < (T)kAlphaThres) continue;
                    const T vis = alpha * Tr;
                    cr += R.r * vis;
                    cg += R.g * vis;
                    cb += R.bl * vis;
                    cur = k;
                    Tr = Tr * (1 - alpha);
                    mg = std::min(mg, rel_margin(Tr, kTThres));
                    if (Tr <= (T)kTThres) break;
                }
                const size_t pix = (size_t)i * W + j;
                rgb[3 * pix + 0] = cr;
                rgb[3 * pix + 1] = cg;
                rgb[3 * pix + 2] = cb;
                Tout[pix] = Tr;
                last[pix] = cur;
                if (margin) margin[pix] = mg;
            }
        }
    }
    if (e_fwd) *e_fwd = E;
}

template <typename T>
inline void atomic_add(T *p, T v) {
#pragma omp atomic
    *p += v;
}

// grad_mode 0: exact dL/dalpha = T_before * c - (colour behind) / (1 - alpha);
// grad_mode 1: the reference's T_before * (c - colour behind)
//              (tile_rasterize_layer.cu:491-495), which drops the 1/T_{k+1}
//              normalisation of the colour behind.
//
// With px_margin (comp_fwd's per-pixel margins) and gs2_amb: for every pixel
// whose forward took a decision within amb_margin of its threshold, gs2_amb
// receives a bound on how much a flip of that decision in fp32 changes each
// gradient term the pixel contributes:
//   - a splat whose own alpha-threshold or clamp decision is within the
//     margin: its whole term (the threshold switches the term on or off, the
//     clamp switches its falloff and conic parts);
//   - every other splat at that pixel: 2 a_f / (1 - a_f) of its term's
//     magnitude, where a_f is the largest alpha of a threshold flipper (a
//     splat of alpha a_f appearing or vanishing scales T behind it by
//     1 - a_f and moves the colour behind the splats in front by at most
//     a_f T c); clamp flips move alpha by rounding only;
//   - a termination (T <= 1e-4) within the margin: the whole term of every
//     splat at the pixel, and of the next contributor after the stop, which
//     fp32 may still include.
template <typename T>
struct BwdTerms {
    T v[9];    // fields: 0 x, 1 y, 2 cxx, 3 cxy, 4 cyy, 5 r, 6 g, 7 b, 8 o
    T mag[9];  // magnitude of each term (dL/dalpha's two products summed apart)
};

template <typename T>
inline BwdTerms<T> bwd_terms(const Rec<T> &R, T dx, T dy, T d2, T g, T alpha, T Tb, T ra,
                             T ar, T ag, T ab, T gr, T gg, T gb, int grad_mode) {
    BwdTerms<T> t{};
    const T fac = alpha * Tb;
    t.v[5] = fac * gr;
    t.v[6] = fac * gg;
    t.v[7] = fac * gb;
    for (int f = 5; f < 8; ++f) t.mag[f] = std::fabs(t.v[f]);
    if (R.o * g <= (T)kMaxAlpha) {
        const T va = grad_mode == 1
                         ? Tb * ((R.r - ar) * gr + (R.g - ag) * gg + (R.bl - ab) * gb)
                         : Tb * (R.r * gr + R.g * gg + R.bl * gb) -
                               ra * (ar * gr + ag * gg + ab * gb);
        const T va_mag =
            grad_mode == 1
                ? Tb * ((std::fabs(R.r) + std::fabs(ar)) * std::fabs(gr) +
                        (std::fabs(R.g) + std::fabs(ag)) * std::fabs(gg) +
                        (std::fabs(R.bl) + std::fabs(ab)) * std::fabs(gb))
                : Tb * (std::fabs(R.r * gr) + std::fabs(R.g * gg) + std::fabs(R.bl * gb)) +
                      ra * (std::fabs(ar * gr) + std::fabs(ag * gg) + std::fabs(ab * gb));
        t.v[8] = g * va;
        t.mag[8] = g * va_mag;
        const T vd2 = (T)0.5 * (-R.o * g * va);
        const T vd2_mag = (T)0.5 * R.o * g * va_mag;
        const T ddx = 2 * (dx * R.a + dy * R.b);
        const T ddy = 2 * (dx * R.b + dy * R.c);
        t.v[0] = vd2 * ddx;
        t.v[1] = vd2 * ddy;
        t.v[2] = vd2 * ((T)-0.25 * ddx * ddx);
        t.v[4] = vd2 * ((T)-0.25 * ddy * ddy);
        t.v[3] = vd2 * ((T)-2 * R.b * d2 - 2 * dx * dy * R.idet);
        t.mag[0] = vd2_mag * std::fabs(ddx);
        t.mag[1] = vd2_mag * std::fabs(ddy);
        t.mag[2] = vd2_mag * ((T)0.25 * ddx * ddx);
        t.mag[4] = vd2_mag * ((T)0.25 * ddy * ddy);
        t.mag[3] = vd2_mag * (std::fabs((T)2 * R.b * d2) + std::fabs(2 * dx * dy * R.idet));
    }
    return t;
}

template <typename T>
void comp_bwd(const T *s2, int n, const uint32_t *flat, const int32_t *ranges, int W, int H,
              const T *Tfin, const int32_t *last, const T *dLdC, T *gs2, int64_t *e_bwd,
              int grad_mode, const double *px_margin = nullptr, double amb_margin = 0.0,
              T *gs2_amb = nullptr) {
    const int ntx = (W + 15) / 16, nty = (H + 15) / 16;
    int64_t E = 0;
    const bool want_amb = px_margin && gs2_amb;
    // gs2 fields: 0 x, 1 y, (2 z untouched), 3 cxx, 4 cxy, 5 cyy, 6 r, 7 g, 8 b, 9 o
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : E)
    for (int tile = 0; tile < ntx * nty; ++tile) {
        const int ty = tile / ntx, tx = tile % ntx;
        const int rs = ranges[2 * tile], re = ranges[2 * tile + 1];
        if (re <= rs) continue;
        std::vector<Rec<T>> recs;
        recs.reserve(re - rs);
        for (int k = rs; k < re; ++k) recs.push_back(rec_of<T>(s2, n, flat[k]));
        std::vector<T> acc((size_t)(re - rs) * 9, (T)0);
        std::vector<T> amb(want_amb ? (size_t)(re - rs) * 9 : 0, (T)0);
        for (int ly = 0; ly < 16; ++ly) {
            for (int lx = 0; lx < 16; ++lx) {
                const int i = ty * 16 + ly, j = tx * 16 + lx;
                if (i >= H || j >= W) continue;
                const size_t pix = (size_t)i * W + j;
                const T px = pixel_x(j, W), py = pixel_y(i, H);
                T Tr = Tfin[pix];
                const int32_t lst = last[pix];
                const T gr = dLdC[3 * pix], gg = dLdC[3 * pix + 1], gb = dLdC[3 * pix + 2];
                const bool pamb = want_amb && px_margin[pix] <= amb_margin;
                // at an ambiguous pixel: which kind of decision is near its
                // threshold (a front-to-back pass as in comp_fwd)
                bool term_amb = false;
                T other_scale = 0;
                if (pamb) {
                    T Tf = 1, af = 0;
                    for (int k = rs; k < re; ++k) {
                        const Rec<T> &R = recs[(size_t)(k - rs)];
                        const T dx = R.x - px, dy = R.y - py;
                        const T d2 = dx * dx * R.a + 2 * dx * dy * R.b + dy * dy * R.c;
                        if ((T)0.5 * d2 < 0) continue;
                        const T og = R.o * std::exp((T)-0.5 * d2);
                        const T alpha = std::min((T)kMaxAlpha, og);
                        if (rel_margin(alpha, kAlphaThres) <= amb_margin)
                            af = std::max(af, std::max(alpha, (T)kAlphaThres));
                        if (alpha < (T)kAlphaThres) continue;
                        Tf = Tf * (1 - alpha);
                        if (rel_margin(Tf, kTThres) <= amb_margin) term_amb = true;
                        if (Tf <= (T)kTThres) break;
                    }
                    other_scale = term_amb ? (T)1 : 2 * af / (1 - af);
                }
                T ar = 0, ag = 0, ab = 0;
                const int kstart = std::min(re - 1, (int)lst);
                if (pamb && term_amb && Tr <= (T)kTThres) {
                    // the next contributor after a near-threshold stop
                    for (int k = kstart + 1; k < re; ++k) {
                        const Rec<T> &R = recs[(size_t)(k - rs)];
                        const T dx = R.x - px, dy = R.y - py;
                        const T d2 = dx * dx * R.a + 2 * dx * dy * R.b + dy * dy * R.c;
                        if ((T)0.5 * d2 < 0) continue;
                        const T g = std::exp((T)-0.5 * d2);
                        const T alpha = std::min((T)kMaxAlpha, R.o * g);
                        if (alpha < (T)kAlphaThres) continue;
                        const T ra = (T)1 / std::max((T)1 - alpha, (T)1e-6);
                        const BwdTerms<T> t = bwd_terms(R, dx, dy, d2, g, alpha, Tr, ra, (T)0,
                                                        (T)0, (T)0, gr, gg, gb, grad_mode);
                        T *B = &amb[(size_t)(k - rs) * 9];
                        for (int f = 0; f < 9; ++f) B[f] += t.mag[f];
                        break;
                    }
                }
                for (int k = kstart; k >= rs; --k) {
                    ++E;
                    const Rec<T> &R = recs[(size_t)(k - rs)];
                    const T dx = R.x - px, dy = R.y - py;
                    const T d2 = dx * dx * R.a + 2 * dx * dy * R.b + dy * dy * R.c;
                    const T sigma = (T)0.5 * d2;
                    if (sigma < 0) continue;
                    const T g = std::exp(-sigma);
                    const T alpha = std::min((T)kMaxAlpha, R.o * g);
                    if (alpha < (T)kAlphaThres) {
                        if (pamb && rel_margin(alpha, kAlphaThres) <= amb_margin) {
                            // a splat fp32 may count: its terms at T before it
                            const T ra = (T)1 / ((T)1 - (T)kAlphaThres);
                            const BwdTerms<T> t =
                                bwd_terms(R, dx, dy, d2, g, (T)kAlphaThres, Tr * ra, ra, ar, ag,
                                          ab, gr, gg, gb, grad_mode);
                            T *B = &amb[(size_t)(k - rs) * 9];
                            for (int f = 0; f < 9; ++f) B[f] += t.mag[f];
                        }
                        continue;
                    }
                    const T ra = (T)1 / std::max((T)1 - alpha, (T)1e-6);
                    const T Tb = Tr * ra;
                    const T fac = alpha * Tb;
                    const BwdTerms<T> t =
                        bwd_terms(R, dx, dy, d2, g, alpha, Tb, ra, ar, ag, ab, gr, gg, gb,
                                  grad_mode);
                    T *A = &acc[(size_t)(k - rs) * 9];
                    for (int f = 0; f < 9; ++f) A[f] += t.v[f];
                    if (pamb) {
                        T *B = &amb[(size_t)(k - rs) * 9];
                        const bool own = rel_margin(alpha, kAlphaThres) <= amb_margin ||
                                         rel_margin(R.o * g, kMaxAlpha) <= amb_margin;
                        const T sc = own ? (T)1 : other_scale;
                        for (int f = 0; f < 9; ++f) B[f] += sc * t.mag[f];
                        if (rel_margin(R.o * g, kMaxAlpha) <= amb_margin &&
                            R.o * g > (T)kMaxAlpha) {
                            // clamped, but fp32 may not clamp: the gated terms
                            const BwdTerms<T> u =
                                bwd_terms(R, dx, dy, d2, g * ((T)kMaxAlpha / (R.o * g)), alpha,
                                          Tb, ra, ar, ag, ab, gr, gg, gb, grad_mode);
                            for (int f = 0; f < 9; ++f) B[f] += u.mag[f];
                        }
                    }
                    ar += R.r * fac;
                    ag += R.g * fac;
                    ab += R.bl * fac;
                    Tr = Tb;
                }
            }
        }
        // fields of acc: 0 x, 1 y, 2 cxx, 3 cxy, 4 cyy, 5 r, 6 g, 7 b, 8 o
        static const int dst[9] = {0, 1, 3, 4, 5, 6, 7, 8, 9};
        for (int k = rs; k < re; ++k) {
            const uint32_t gi = flat[k];
            const T *A = &acc[(size_t)(k - rs) * 9];
            for (int f = 0; f < 9; ++f)
                if (A[f] != 0) atomic_add(&gs2[(size_t)dst[f] * n + gi], A[f]);
            if (want_amb) {
                const T *B = &amb[(size_t)(k - rs) * 9];
                for (int f = 0; f < 9; ++f)
                    if (B[f] != 0) atomic_add(&gs2_amb[(size_t)dst[f] * n + gi], B[f]);
            }
        }
    }
    if (e_bwd) *e_bwd = E;
}

// ---------------------------------------------------------------- sort / ranges
// Pairs of visible splats emitted in depth order (stable by splat index),
// tiles row-major, first `cap` kept; then stably sorted by the 64-bit key.
// Without the cap this equals the reference's emit-then-SortPairs output.
int64_t assign_and_sort(const float *s2, int n, int W, int H, int64_t cap, uint64_t *keys,
                        uint32_t *ids, int64_t *wanted, uint32_t *visible, int32_t *rects,
                        uint32_t *counts) {
    const int ntx = (W + 15) / 16, nty = (H + 15) / 16;
    std::vector<Rect> R((size_t)n);
    std::vector<uint32_t> dk((size_t)n);
    uint32_t vis = 0;
#pragma omp parallel for schedule(static) reduction(+ : vis)
    for (int i = 0; i < n; ++i) {
        R[i] = tile_rect(s2, n, i, W, H, ntx, nty);
        dk[i] = depth_bits(s2[2 * (size_t)n + i]);
        vis += R[i].vis ? 1u : 0u;
    }
    std::vector<uint32_t> order((size_t)n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return dk[a] < dk[b]; });
    int64_t total = 0;
    for (int i = 0; i < n; ++i)
        if (R[i].vis) total += (int64_t)(R[i].tx1 - R[i].tx0 + 1) * (R[i].ty1 - R[i].ty0 + 1);
    const int64_t kept = std::min(total, cap);
    // emit in depth order
    std::vector<uint32_t> tl((size_t)kept), gl((size_t)kept);
    int64_t slot = 0;
    for (int k = 0; k < n && slot < kept; ++k) {
        const uint32_t g = order[k];
        const Rect &r = R[g];
        if (!r.vis) continue;
        for (int ty = r.ty0; ty <= r.ty1 && slot < kept; ++ty)
            for (int tx = r.tx0; tx <= r.tx1 && slot < kept; ++tx) {
                tl[slot] = (uint32_t)(ty * ntx + tx);
                gl[slot] = g;
                ++slot;
            }
    }
    // stable counting sort by tile
    const int ntiles = ntx * nty;
    std::vector<int64_t> cnt((size_t)ntiles + 1, 0);
    for (int64_t s = 0; s < kept; ++s) ++cnt[tl[s] + 1];
    for (int t = 0; t < ntiles; ++t) cnt[t + 1] += cnt[t];
    for (int64_t s = 0; s < kept; ++s) {
        const int64_t p = cnt[tl[s]]++;
        keys[p] = ((uint64_t)tl[s] << 32) | dk[gl[s]];
        ids[p] = gl[s];
    }
    if (wanted) *wanted = total;
    if (visible) *visible = vis;
    if (rects || counts) {
        for (int i = 0; i < n; ++i) {
            if (rects) {
                rects[4 * i + 0] = R[i].vis ? R[i].tx0 : 0;
                rects[4 * i + 1] = R[i].vis ? R[i].tx1 : -1;
                rects[4 * i + 2] = R[i].vis ? R[i].ty0 : 0;
                rects[4 * i + 3] = R[i].vis ? R[i].ty1 : -1;
            }
            if (counts)
                counts[i] = R[i].vis ? (uint32_t)((R[i].tx1 - R[i].tx0 + 1) *
                                                  (R[i].ty1 - R[i].ty0 + 1))
                                     : 0u;
        }
    }
    return kept;
}

void tile_ranges(const uint64_t *keys, int64_t n, int ntiles, int32_t *ranges) {
    std::fill(ranges, ranges + 2 * (size_t)ntiles, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int t = (int)(keys[i] >> 32);
        if (t < 0 || t >= ntiles) continue;
        if (i == 0 || (int)(keys[i - 1] >> 32) != t) ranges[2 * t] = (int32_t)i;
        if (i == n - 1 || (int)(keys[i + 1] >> 32) != t) ranges[2 * t + 1] = (int32_t)(i + 1);
    }
}

// ---------------------------------------------------------------- loss / Adam
template <typename T>
T mse(const T *rgb, const T *tgt, int64_t npix, T *grad) {
    const T scale = (T)1 / (T)npix;
    T loss = 0;
#pragma omp parallel for schedule(static) reduction(+ : loss)
    for (int64_t p = 0; p < npix; ++p) {
        T d[3];
        for (int c = 0; c < 3; ++c) d[c] = rgb[3 * p + c] - tgt[3 * p + c];
        loss += (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * scale;
        if (grad)
            for (int c = 0; c < 3; ++c) grad[3 * p + c] = (T)2 * scale * d[c];
    }
    return loss;
}

template <typename T>
void adam(T *p, const T *g, T *m, T *v, int64_t n, T lr, T b1, T b2, T eps, int step) {
    const T bc1 = (T)1 / ((T)1 - std::pow(b1, (T)step));
    const T bc2 = (T)1 / ((T)1 - std::pow(b2, (T)step));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const T gi = g[i];
        m[i] = b1 * m[i] + (1 - b1) * gi;
        v[i] = b2 * v[i] + (1 - b2) * gi * gi;
        p[i] -= lr * (m[i] * bc1) / (std::sqrt(v[i] * bc2) + eps);
    }
}

// ---------------------------------------------------------------- full fp32 step
struct Fit {
    int n, W, H, deg, nb;
    float pv[16], cam[3], lr[6], b1, b2, eps;
    int64_t cap;
    int step_no = 0;
    int grad_mode = 0;
    std::vector<float> params, m, v, grads, s3, gs3, s2, gs2, rgb, Tf, dLdC, target;
    std::vector<int32_t> last, ranges;
    std::vector<uint64_t> keys;
    std::vector<uint32_t> ids;
    float loss = 0;
};

// ---------------------------------------------------------------- densify
// Adaptive density control (Kerbl et al., SIGGRAPH 2023, sec. 5.2; a TODO in
// the reference, README.md:10): the engine's rules, restated. Decisions are
// float comparisons on the raw parameters (as the engine takes them); the
// split samples are computed in double from the same hashed uniforms.
inline uint64_t dens_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline double dens_uniform(uint64_t seed, uint32_t i, uint32_t s, uint32_t k) {
    const uint64_t h = dens_mix64(seed ^ dens_mix64(((uint64_t)i << 8) | (s << 2) | k));
    return (double)((uint32_t)(h >> 40) + 1u) / 16777216.0;
}

enum { DENS_PRUNE = 0, DENS_KEEP = 1, DENS_CLONE = 2, DENS_SPLIT = 3 };

int64_t densify(const float *slab, const float *m, const float *v, int n, int nfields,
                const float *accum, const uint32_t *visits, float grad_threshold,
                float clone_max_log_scale, float min_logit, float prune_max_log_scale,
                int prune_big, float split_log_div, uint64_t seed, uint8_t *action,
                double *out, float *out_m, float *out_v) {
    std::vector<uint32_t> keep((size_t)n), added((size_t)n);
    for (int i = 0; i < n; ++i) {
        const uint32_t nv = visits[i];
        const float g = nv ? accum[i] / (float)nv : 0.f;
        const float ls = std::max(std::max(slab[3 * (size_t)n + i], slab[4 * (size_t)n + i]),
                                  slab[5 * (size_t)n + i]);
        const bool faint = slab[13 * (size_t)n + (size_t)(nfields - 14) * n + i] < min_logit;
        const bool hot = g >= grad_threshold;
        const bool small = ls <= clone_max_log_scale;
        const bool too_big = prune_big && ls > prune_max_log_scale;
        const bool child_too_big = prune_big && ls - split_log_div > prune_max_log_scale;
        int a;
        uint32_t k = 0, mm = 0;
        if (faint) {
            a = DENS_PRUNE;
        } else if (hot && small) {
            a = DENS_CLONE;
            k = mm = too_big ? 0u : 1u;
        } else if (hot) {
            a = DENS_SPLIT;
            mm = child_too_big ? 0u : 2u;
        } else {
            a = too_big ? DENS_PRUNE : DENS_KEEP;
            k = too_big ? 0u : 1u;
        }
        action[i] = (uint8_t)a;
        keep[i] = k;
        added[i] = mm;
    }
    uint64_t nkeep = 0, nadd = 0;
    for (int i = 0; i < n; ++i) {
        nkeep += keep[i];
        nadd += added[i];
    }
    const size_t n_out = nkeep + nadd;
    if (!out) return (int64_t)n_out;
    size_t jk = 0, ja = nkeep;
    for (int i = 0; i < n; ++i) {
        if (keep[i]) {
            for (int f = 0; f < nfields; ++f) {
                out[f * n_out + jk] = slab[(size_t)f * n + i];
                out_m[f * n_out + jk] = m[(size_t)f * n + i];
                out_v[f * n_out + jk] = v[(size_t)f * n + i];
            }
            ++jk;
        }
        for (uint32_t s = 0; s < added[i]; ++s, ++ja) {
            for (int f = 0; f < nfields; ++f) {
                out[f * n_out + ja] = slab[(size_t)f * n + i];
                out_m[f * n_out + ja] = 0.f;
                out_v[f * n_out + ja] = 0.f;
            }
            if (action[i] != DENS_SPLIT) continue;
            const double u0 = dens_uniform(seed, (uint32_t)i, s, 0);
            const double u1 = dens_uniform(seed, (uint32_t)i, s, 1);
            const double u2 = dens_uniform(seed, (uint32_t)i, s, 2);
            const double u3 = dens_uniform(seed, (uint32_t)i, s, 3);
            const double r0 = std::sqrt(-2.0 * std::log(u0)), r1 = std::sqrt(-2.0 * std::log(u2));
            const double two_pi = 6.283185307179586;
            const double z[3] = {r0 * std::cos(two_pi * u1), r0 * std::sin(two_pi * u1),
                                 r1 * std::cos(two_pi * u3)};
            const RSq<double> rs = make_rs<double>(
                slab[3 * (size_t)n + i], slab[4 * (size_t)n + i], slab[5 * (size_t)n + i],
                slab[6 * (size_t)n + i], slab[7 * (size_t)n + i], slab[8 * (size_t)n + i],
                slab[9 * (size_t)n + i]);
            for (int r = 0; r < 3; ++r) {
                double d = 0;
                for (int c = 0; c < 3; ++c) d += rs.M[r][c] * z[c];
                out[r * n_out + ja] = (double)slab[(size_t)r * n + i] + d;
            }
            for (int f = 3; f < 6; ++f)
                out[f * n_out + ja] = (double)(slab[(size_t)f * n + i] - split_log_div);
        }
    }
    return (int64_t)n_out;
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

// Densification (see densify above); with out == NULL only the actions and
// the new count.
int64_t or_densify(const float *slab, const float *m, const float *v, int n, int nfields,
                   const float *accum, const uint32_t *visits, float grad_threshold,
                   float clone_max_log_scale, float min_logit, float prune_max_log_scale,
                   int prune_big, float split_log_div, uint64_t seed, uint8_t *action,
                   double *out, float *out_m, float *out_v) {
    return densify(slab, m, v, n, nfields, accum, visits, grad_threshold, clone_max_log_scale,
                   min_logit, prune_max_log_scale, prune_big, split_log_div, seed, action, out,
                   out_m, out_v);
}

#define OR_DEFINE(SUF, T)                                                                     \
    void or_act_fwd_##SUF(const T *params, int n, int deg, const T *cam, T *s3) {             \
        act_fwd<T>(params, n, deg, cam, s3);                                                  \
    }                                                                                         \
    void or_act_bwd_##SUF(const T *params, const T *s3, const T *gs3, int n, int deg,         \
                          const T *cam, T *gparams, int grad_mode) {                          \
        act_bwd<T>(params, s3, gs3, n, deg, cam, gparams, grad_mode);                         \
    }                                                                                         \
    void or_proj_fwd_##SUF(const T *s3, int n, const T *pv, T *s2) {                          \
        proj_fwd<T>(s3, n, pv, s2);                                                           \
    }                                                                                         \
    void or_proj_bwd_##SUF(const T *s3, const T *gs2, int n, const T *pv, T *gs3,             \
                           int grad_mode) {                                                   \
        proj_bwd<T>(s3, gs2, n, pv, gs3, grad_mode);                                          \
    }                                                                                         \
    void or_comp_fwd_##SUF(const T *s2, int n, const uint32_t *flat, const int32_t *ranges,   \
                           int W, int H, T *rgb, T *Tout, int32_t *last, double *margin,      \
                           int64_t *e_fwd) {                                                  \
        comp_fwd<T>(s2, n, flat, ranges, W, H, rgb, Tout, last, margin, e_fwd);               \
    }                                                                                         \
    void or_comp_bwd_##SUF(const T *s2, int n, const uint32_t *flat, const int32_t *ranges,   \
                           int W, int H, const T *Tfin, const int32_t *last, const T *dLdC,   \
                           T *gs2, int64_t *e_bwd, int grad_mode, const double *px_margin,    \
                           double amb_margin, T *gs2_amb) {                                   \
        comp_bwd<T>(s2, n, flat, ranges, W, H, Tfin, last, dLdC, gs2, e_bwd, grad_mode,       \
                    px_margin, amb_margin, gs2_amb);                                          \
    }                                                                                         \
    T or_mse_##SUF(const T *rgb, const T *tgt, int64_t npix, T *grad) {                       \
        return mse<T>(rgb, tgt, npix, grad);                                                  \
    }                                                                                         \
    void or_adam_##SUF(T *p, const T *g, T *m, T *v, int64_t n, T lr, T b1, T b2, T eps,      \
                       int step) {                                                            \
        adam<T>(p, g, m, v, n, lr, b1, b2, eps, step);                                        \
    }

OR_DEFINE(f32, float)
OR_DEFINE(f64, double)

int64_t or_tile_assign_sort(const float *s2, int n, int W, int H, int64_t cap, uint64_t *keys,
                            uint32_t *ids, int64_t *wanted, uint32_t *visible, int32_t *rects,
                            uint32_t *counts) {
    return assign_and_sort(s2, n, W, H, cap, keys, ids, wanted, visible, rects, counts);
}

void or_tile_ranges(const uint64_t *keys, int64_t n, int ntiles, int32_t *ranges) {
    tile_ranges(keys, n, ntiles, ranges);
}

uint64_t or_make_key(uint32_t tile, float depth) {
    return ((uint64_t)tile << 32) | depth_bits(depth);
}

int or_set_threads(int t) {
    if (t > 0) omp_set_num_threads(t);
    return omp_get_max_threads();
}

// fp32 fit (the CPU baseline): one object holds the whole state.
void *or_fit_create(int n, int W, int H, int deg, const float *pv, const float *cam,
                    const float *lr6, float b1, float b2, float eps, int64_t cap,
                    const float *params, const float *target, int grad_mode) {
    Fit *f = new Fit();
    f->grad_mode = grad_mode;
    f->n = n;
    f->W = W;
    f->H = H;
    f->deg = deg;
    f->nb = bands_of(deg);
    std::memcpy(f->pv, pv, sizeof(f->pv));
    std::memcpy(f->cam, cam, sizeof(f->cam));
    std::memcpy(f->lr, lr6, sizeof(f->lr));
    f->b1 = b1;
    f->b2 = b2;
    f->eps = eps;
    f->cap = cap > 0 ? cap : ((int64_t)1 << 25);
    const size_t np = (size_t)(14 + 3 * f->nb) * n;
    f->params.assign(params, params + np);
    f->m.assign(np, 0.f);
    f->v.assign(np, 0.f);
    f->grads.assign(np, 0.f);
    f->s3.assign((size_t)13 * n, 0.f);
    f->gs3.assign((size_t)13 * n, 0.f);
    f->s2.assign((size_t)10 * n, 0.f);
    f->gs2.assign((size_t)10 * n, 0.f);
    const size_t npix = (size_t)W * H;
    f->rgb.assign(3 * npix, 0.f);
    f->dLdC.assign(3 * npix, 0.f);
    f->Tf.assign(npix, 1.f);
    f->last.assign(npix, 0);
    f->target.assign(target, target + 3 * npix);
    f->ranges.assign(2 * (size_t)((W + 15) / 16) * ((H + 15) / 16), 0);
    return f;
}

void or_fit_destroy(void *h) { delete static_cast<Fit *>(h); }

// One ImageFitter::step: zeroGrad, forward, backward, Adam. Returns the loss
// of this step's forward.
float or_fit_step(void *h) {
    Fit &f = *static_cast<Fit *>(h);
    const int n = f.n;
    act_fwd<float>(f.params.data(), n, f.deg, f.cam, f.s3.data());
    proj_fwd<float>(f.s3.data(), n, f.pv, f.s2.data());
    int64_t wanted = 0;
    uint32_t vis = 0;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {  // upper bound for the key buffers
        Rect r = tile_rect(f.s2.data(), n, i, f.W, f.H, (f.W + 15) / 16, (f.H + 15) / 16);
        if (r.vis) total += (int64_t)(r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1);
    }
    const int64_t want = std::min(total, f.cap);
    if ((int64_t)f.keys.size() < want) {
        f.keys.resize((size_t)want);
        f.ids.resize((size_t)want);
    }
    const int64_t ni = assign_and_sort(f.s2.data(), n, f.W, f.H, f.cap, f.keys.data(),
                                       f.ids.data(), &wanted, &vis, nullptr, nullptr);
    const int ntiles = ((f.W + 15) / 16) * ((f.H + 15) / 16);
    tile_ranges(f.keys.data(), ni, ntiles, f.ranges.data());
    comp_fwd<float>(f.s2.data(), n, f.ids.data(), f.ranges.data(), f.W, f.H, f.rgb.data(),
                    f.Tf.data(), f.last.data(), nullptr, nullptr);
    const int64_t npix = (int64_t)f.W * f.H;
    f.loss = mse<float>(f.rgb.data(), f.target.data(), npix, f.dLdC.data());
    std::fill(f.gs2.begin(), f.gs2.end(), 0.f);
    comp_bwd<float>(f.s2.data(), n, f.ids.data(), f.ranges.data(), f.W, f.H, f.Tf.data(),
                    f.last.data(), f.dLdC.data(), f.gs2.data(), nullptr, f.grad_mode);
    proj_bwd<float>(f.s3.data(), f.gs2.data(), n, f.pv, f.gs3.data(), f.grad_mode);
    act_bwd<float>(f.params.data(), f.s3.data(), f.gs3.data(), n, f.deg, f.cam, f.grads.data(),
                   f.grad_mode);
    ++f.step_no;
    // groups: pos(3) lr0, scale(3) lr1, rot(4) lr2, dc(3) lr3, rest(3*nb) lr4, logit lr5
    const int64_t N = n;
    auto run = [&](int64_t off, int64_t cnt, float lr) {
        adam<float>(f.params.data() + off, f.grads.data() + off, f.m.data() + off,
                    f.v.data() + off, cnt, lr, f.b1, f.b2, f.eps, f.step_no);
    };
    run(0, 3 * N, f.lr[0]);
    run(3 * N, 3 * N, f.lr[1]);
    run(6 * N, 4 * N, f.lr[2]);
    run(10 * N, 3 * N, f.lr[3]);
    if (f.nb) run(13 * N, 3 * f.nb * N, f.lr[4]);
    run((13 + 3 * (int64_t)f.nb) * N, N, f.lr[5]);
    return f.loss;
}

void or_fit_get(void *h, float *params, float *rgb) {
    Fit &f = *static_cast<Fit *>(h);
    if (params) std::memcpy(params, f.params.data(), f.params.size() * sizeof(float));
    if (rgb) std::memcpy(rgb, f.rgb.data(), f.rgb.size() * sizeof(float));
}

}  // extern "C"
