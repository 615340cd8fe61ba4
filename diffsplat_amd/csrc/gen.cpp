// Deterministic synthetic workloads and OpenGL-convention camera matrices
// (host only). The random stream is splitmix64 keyed by the seed, so every
// machine generates the same inputs.
//
// Camera matrices follow GLM's defaults as used by the reference
// (glm::ortho / glm::perspective / glm::lookAt at src/pipelines/image_fitter.cpp:26-29,
// src/camera/arcball_camera.cpp:136-144): right-handed, clip z in [-1, 1],
// column-major storage m[col * 4 + row].
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ds_api.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kSH_C0 = 0.28209479177387814f;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

// Slab field pointers for count splats with nb SH bands.
struct Slab {
    float *f[13];
    float *rest_r, *rest_g, *rest_b, *logit;
    Slab(float *p, int n, int nb) {
        for (int k = 0; k < 13; ++k) f[k] = p + (size_t)k * n;
        rest_r = p + 13 * (size_t)n;
        rest_g = rest_r + (size_t)nb * n;
        rest_b = rest_g + (size_t)nb * n;
        logit = rest_b + (size_t)nb * n;
    }
};

void random_unit_quat(Rng &r, float q[4]) {
    double a[4], n2 = 0;
    for (int k = 0; k < 4; ++k) {
        a[k] = r.normal();
        n2 += a[k] * a[k];
    }
    if (n2 < 1e-24) {
        q[0] = 1.f;
        q[1] = q[2] = q[3] = 0.f;
        return;
    }
    const double inv = 1.0 / std::sqrt(n2);
    for (int k = 0; k < 4; ++k) q[k] = (float)(a[k] * inv);
}

}  // namespace

extern "C" {

int ds_sh_bands(int d) { return d <= 0 ? 0 : d == 1 ? 3 : d == 2 ? 8 : 15; }

int64_t ds_slab_floats(int count, int sh_degree) {
    return (int64_t)(14 + 3 * ds_sh_bands(sh_degree)) * (int64_t)count;
}

void ds_gen_target_image(uint64_t seed, int W, int H, int n_ellipses, float *out) {
    Rng r(seed);
    double c0[3], gx[3], gy[3];
    for (int c = 0; c < 3; ++c) {
        c0[c] = r.uniform(0.25, 0.75);
        gx[c] = r.uniform(-0.3, 0.3);
        gy[c] = r.uniform(-0.3, 0.3);
    }
    struct Sine {
        double amp[3], fx, fy, ph;
    } sn[3];
    for (auto &s : sn) {
        for (double &a : s.amp) a = r.uniform(0.04, 0.12);
        s.fx = r.uniform(-6.0, 6.0);
        s.fy = r.uniform(-6.0, 6.0);
        s.ph = r.uniform(0.0, 2.0 * kPi);
    }
    struct Ell {
        double cx, cy, ra, rb, ca, sa, col[3];
    };
    std::vector<Ell> el((size_t)(n_ellipses > 0 ? n_ellipses : 0));
    const double m = (double)(W < H ? W : H);
    for (auto &e : el) {
        e.cx = r.uniform(0.0, (double)W);
        e.cy = r.uniform(0.0, (double)H);
        e.ra = r.uniform(0.02, 0.15) * m;
        e.rb = r.uniform(0.02, 0.15) * m;
        const double ang = r.uniform(0.0, kPi);
        e.ca = std::cos(ang);
        e.sa = std::sin(ang);
        for (double &c : e.col) c = r.uniform();
    }
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const double u = (x + 0.5) / W, v = (y + 0.5) / H;
            double px[3];
            for (int c = 0; c < 3; ++c) px[c] = c0[c] + gx[c] * (u - 0.5) + gy[c] * (v - 0.5);
            for (const auto &s : sn) {
                const double w = std::sin(2.0 * kPi * (s.fx * u + s.fy * v) + s.ph);
                for (int c = 0; c < 3; ++c) px[c] += s.amp[c] * w;
            }
            for (double &c : px) c = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
            const double fx = x + 0.5, fy = y + 0.5;
            for (const auto &e : el) {
                const double dx = fx - e.cx, dy = fy - e.cy;
                const double a = (dx * e.ca + dy * e.sa) / e.ra;
                const double b = (-dx * e.sa + dy * e.ca) / e.rb;
                if (a * a + b * b <= 1.0)
                    for (int c = 0; c < 3; ++c) px[c] = e.col[c];
            }
            float *o = out + 3 * ((size_t)y * W + x);
            for (int c = 0; c < 3; ++c) o[c] = (float)px[c];
        }
    }
}

void ds_gen_frustum_splats(uint64_t seed, int count, float fovy_deg, float aspect, float zmin,
                           float zmax, float ls_mean, float ls_std, float logit_opacity,
                           float *slab) {
    Rng r(seed ^ 0x5EED5EED5EEDull);
    Slab s(slab, count, 0);
    const double th = std::tan(0.5 * fovy_deg * kPi / 180.0);
    for (int i = 0; i < count; ++i) {
        const double d = r.uniform(zmin, zmax);
        const double xn = r.uniform(-1.0, 1.0), yn = r.uniform(-1.0, 1.0);
        s.f[0][i] = (float)(xn * d * th * aspect);
        s.f[1][i] = (float)(yn * d * th);
        s.f[2][i] = (float)(-d);
        for (int k = 0; k < 3; ++k) s.f[3 + k][i] = (float)(ls_mean + ls_std * r.normal());
        float q[4];
        random_unit_quat(r, q);
        for (int k = 0; k < 4; ++k) s.f[6 + k][i] = q[k];
        for (int k = 0; k < 3; ++k) s.f[10 + k][i] = (float)((r.uniform() - 0.5) / kSH_C0);
        s.logit[i] = logit_opacity;
    }
}

void ds_gen_cluster_scene(uint64_t seed, int count, int sh_degree, int n_clusters,
                          float ls_mean, float ls_std, float *slab) {
    Rng r(seed ^ 0xC1A57E5ull);
    const int nb = ds_sh_bands(sh_degree);
    Slab s(slab, count, nb);
    if (n_clusters < 1) n_clusters = 1;
    struct Cl {
        double c[3], R[3][3], sd[3];
    };
    std::vector<Cl> cl((size_t)n_clusters);
    for (auto &k : cl) {
        for (double &v : k.c) v = r.uniform(-0.35, 0.35);
        float q[4];
        random_unit_quat(r, q);
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        std::memcpy(k.R, R, sizeof(R));
        for (double &v : k.sd) v = r.uniform(0.01, 0.12);
    }
    for (int i = 0; i < count; ++i) {
        const Cl &k = cl[(size_t)(r.next() % (uint64_t)n_clusters)];
        double e[3];
        for (int a = 0; a < 3; ++a) e[a] = k.sd[a] * r.normal();
        for (int a = 0; a < 3; ++a) {
            double p = k.c[a] + k.R[a][0] * e[0] + k.R[a][1] * e[1] + k.R[a][2] * e[2];
            p = p < -0.5 ? -0.5 : (p > 0.5 ? 0.5 : p);
            s.f[a][i] = (float)p;
        }
        for (int a = 0; a < 3; ++a) s.f[3 + a][i] = (float)(ls_mean + ls_std * r.normal());
        float q[4];
        random_unit_quat(r, q);
        for (int a = 0; a < 4; ++a) s.f[6 + a][i] = q[a];
        for (int c = 0; c < 3; ++c) s.f[10 + c][i] = (float)((0.2 * r.normal()) / kSH_C0);
        for (int b = 0; b < nb; ++b) {
            s.rest_r[(size_t)b * count + i] = (float)(0.2 * r.normal());
            s.rest_g[(size_t)b * count + i] = (float)(0.2 * r.normal());
            s.rest_b[(size_t)b * count + i] = (float)(0.2 * r.normal());
        }
        s.logit[i] = (float)(2.0 * r.normal());
    }
}

void ds_gen_fitter_init(uint64_t seed, int count, int width, int height, float *slab) {
    Rng r(seed ^ 0xF17E5ull);
    Slab s(slab, count, 0);
    const float hw = width * 0.5f, hh = height * 0.5f;
    const float log_sigma = std::log(2.f);
    for (int i = 0; i < count; ++i) {
        s.f[0][i] = (float)r.uniform(-1.0, 1.0) * hw;
        s.f[1][i] = (float)r.uniform(-1.0, 1.0) * hh;
        s.f[2][i] = (float)r.uniform(-1.0, 1.0) * (hw < hh ? hw : hh) * 0.25f;
        const float sc = (float)r.uniform(-1.0, 1.0) * 0.5f;
        for (int k = 0; k < 3; ++k) s.f[3 + k][i] = log_sigma + sc;
        s.f[6][i] = 1.f;
        s.f[7][i] = s.f[8][i] = s.f[9][i] = 0.f;
        for (int c = 0; c < 3; ++c) s.f[10 + c][i] = ((float)r.uniform() - 0.5f) / kSH_C0;
        const float o = 0.6f + 0.4f * (float)r.uniform();
        s.logit[i] = std::log(o / (1.f - o));
    }
}

void ds_mat_perspective(float fovy, float aspect, float zn, float zf, float m[16]) {
    std::memset(m, 0, 16 * sizeof(float));
    const float t = std::tan(fovy * 0.5f);
    m[0] = 1.f / (aspect * t);
    m[5] = 1.f / t;
    m[10] = -(zf + zn) / (zf - zn);
    m[11] = -1.f;
    m[14] = -(2.f * zf * zn) / (zf - zn);
}

void ds_mat_ortho(float l, float r, float b, float t, float n, float f, float m[16]) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = 2.f / (r - l);
    m[5] = 2.f / (t - b);
    m[10] = -2.f / (f - n);
    m[12] = -(r + l) / (r - l);
    m[13] = -(t + b) / (t - b);
    m[14] = -(f + n) / (f - n);
    m[15] = 1.f;
}

void ds_mat_look_at(const float eye[3], const float ctr[3], const float up[3], float m[16]) {
    float f[3] = {ctr[0] - eye[0], ctr[1] - eye[1], ctr[2] - eye[2]};
    float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float &v : f) v /= fl;
    float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2],
                  f[0] * up[1] - f[1] * up[0]};
    float sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (float &v : s) v /= sl;
    const float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2],
                        s[0] * f[1] - s[1] * f[0]};
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = s[0]; m[4] = s[1]; m[8] = s[2];
    m[1] = u[0]; m[5] = u[1]; m[9] = u[2];
    m[2] = -f[0]; m[6] = -f[1]; m[10] = -f[2];
    m[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    m[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    m[14] = (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]);
    m[15] = 1.f;
}

void ds_mat_mul(const float a[16], const float b[16], float out[16]) {
    float t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float acc = 0.f;
            for (int k = 0; k < 4; ++k) acc += a[k * 4 + r] * b[c * 4 + k];
            t[c * 4 + r] = acc;
        }
    std::memcpy(out, t, sizeof(t));
}

}  // extern "C"
