// Host-side device buffers and SoA containers with the reference's names.
//
//   CudaBuffer<T>        src/cuda/cuda_buffer.h:7-76      -> DeviceBuffer<T>
//   GpuSoA               src/types/gpu_soa.h:21-35
//   Gaussian3DParams/Grads  src/types/gaussian3d.h:34-120
//   Splat3DParams/Grads  src/types/splat3d.h:14-51
//   Splat2DParams/Grads  src/types/splat2d.h:13-46
//
// Difference from the reference: every SoA type owns ONE contiguous slab
// (field-major, count floats per field) and its fields are views into it, so
// a whole parameter set zeroes, copies or runs through Adam as one range.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ds_api.h"

namespace dsh {

inline void check(ds_status st, const char *what) {
    if (st != DS_OK)
        throw std::runtime_error(std::string("[diffsplat] ") + what + ": " +
                                 ds_status_string(st) + ": " + ds_last_error());
}

template <typename T>
class DeviceBuffer {
public:
    T *ptr = nullptr;
    size_t size = 0;

    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), size(o.size) {
        o.ptr = nullptr;
        o.size = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            size = o.size;
            o.ptr = nullptr;
            o.size = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void allocate(size_t n) {
        if (n == size && ptr) return;
        release();
        if (n == 0) return;
        void *p = nullptr;
        check(ds_malloc(&p, n * sizeof(T)), "DeviceBuffer::allocate");
        ptr = static_cast<T *>(p);
        size = n;
    }
    void release() {
        if (ptr) ds_free(ptr);
        ptr = nullptr;
        size = 0;
    }
    // Zero with a fill kernel (safe inside a captured graph).
    void zero(hipStream_t s = nullptr) {
        if (ptr && size)
            check(ds_fill_u32(ptr, 0u, size * sizeof(T) / 4, s), "DeviceBuffer::zero");
    }
    void upload(const T *host, size_t n) {
        check(ds_memcpy_htod(ptr, host, n * sizeof(T)), "DeviceBuffer::upload");
    }
    void download(T *host, size_t n) const {
        check(ds_memcpy_dtoh(host, ptr, n * sizeof(T)), "DeviceBuffer::download");
    }
    operator T *() const { return ptr; }
};

// A set of `nfields` float arrays of `count` floats in one slab.
struct GpuSoA {
    int count = 0;
    DeviceBuffer<float> slab;
    virtual ~GpuSoA() = default;
    virtual int nfields() const = 0;
    virtual void bind() = 0;  // point the named fields into the slab
    virtual void allocate(int n) {
        count = n;
        slab.allocate((size_t)nfields() * (size_t)n);
        bind();
    }
    void zero(hipStream_t s = nullptr) { slab.zero(s); }
    float *field(int k) const { return slab.ptr ? slab.ptr + (size_t)k * count : nullptr; }
};

struct Gaussian3DParams : GpuSoA {
    float *pos_x = nullptr, *pos_y = nullptr, *pos_z = nullptr;
    float *scale_x = nullptr, *scale_y = nullptr, *scale_z = nullptr;
    float *rot_w = nullptr, *rot_x = nullptr, *rot_y = nullptr, *rot_z = nullptr;
    float *sh_dc_r = nullptr, *sh_dc_g = nullptr, *sh_dc_b = nullptr;
    float *sh_rest_r = nullptr, *sh_rest_g = nullptr, *sh_rest_b = nullptr;
    float *logit_opacity = nullptr;
    int sh_num_bands = 0;

    int nfields() const override { return 14 + 3 * sh_num_bands; }
    void allocate(int n) override { allocate(n, 0); }
    void allocate(int n, int sh_degree) {
        sh_num_bands = ds_sh_bands(sh_degree);
        GpuSoA::allocate(n);
    }
    void bind() override {
        float *f[13];
        for (int k = 0; k < 13; ++k) f[k] = field(k);
        pos_x = f[0]; pos_y = f[1]; pos_z = f[2];
        scale_x = f[3]; scale_y = f[4]; scale_z = f[5];
        rot_w = f[6]; rot_x = f[7]; rot_y = f[8]; rot_z = f[9];
        sh_dc_r = f[10]; sh_dc_g = f[11]; sh_dc_b = f[12];
        const size_t nb = (size_t)sh_num_bands * count;
        float *rest = slab.ptr ? slab.ptr + 13 * (size_t)count : nullptr;
        sh_rest_r = sh_num_bands ? rest : nullptr;
        sh_rest_g = sh_num_bands ? rest + nb : nullptr;
        sh_rest_b = sh_num_bands ? rest + 2 * nb : nullptr;
        logit_opacity = slab.ptr ? rest + 3 * nb : nullptr;
    }
    ds_gauss3d view() const {
        return ds_gauss3d{pos_x,   pos_y,   pos_z,     scale_x,   scale_y,   scale_z,
                          rot_w,   rot_x,   rot_y,     rot_z,     sh_dc_r,   sh_dc_g,
                          sh_dc_b, sh_rest_r, sh_rest_g, sh_rest_b, logit_opacity, count,
                          sh_num_bands};
    }
};
using Gaussian3DGrads = Gaussian3DParams;

struct Splat3DParams : GpuSoA {
    float *pos_x = nullptr, *pos_y = nullptr, *pos_z = nullptr;
    float *cov_xx = nullptr, *cov_xy = nullptr, *cov_xz = nullptr;
    float *cov_yy = nullptr, *cov_yz = nullptr, *cov_zz = nullptr;
    float *color_r = nullptr, *color_g = nullptr, *color_b = nullptr;
    float *opacity = nullptr;
    int nfields() const override { return 13; }
    void bind() override {
        float **f[13] = {&pos_x,  &pos_y,  &pos_z,   &cov_xx,  &cov_xy,  &cov_xz, &cov_yy,
                         &cov_yz, &cov_zz, &color_r, &color_g, &color_b, &opacity};
        for (int k = 0; k < 13; ++k) *f[k] = field(k);
    }
    ds_splat3d view() const {
        return ds_splat3d{pos_x,  pos_y,  pos_z,   cov_xx,  cov_xy,  cov_xz,  cov_yy,
                          cov_yz, cov_zz, color_r, color_g, color_b, opacity, count};
    }
};
using Splat3DGrads = Splat3DParams;

struct Splat2DParams : GpuSoA {
    float *pos_x = nullptr, *pos_y = nullptr, *pos_z = nullptr;
    float *cov_xx = nullptr, *cov_xy = nullptr, *cov_yy = nullptr;
    float *color_r = nullptr, *color_g = nullptr, *color_b = nullptr;
    float *opacity = nullptr;
    int nfields() const override { return 10; }
    void bind() override {
        float **f[10] = {&pos_x,  &pos_y,   &pos_z,   &cov_xx,  &cov_xy,
                         &cov_yy, &color_r, &color_g, &color_b, &opacity};
        for (int k = 0; k < 10; ++k) *f[k] = field(k);
    }
    ds_splat2d view() const {
        return ds_splat2d{pos_x,   pos_y,   pos_z,   cov_xx,  cov_xy,
                          cov_yy,  color_r, color_g, color_b, opacity, count};
    }
};
using Splat2DGrads = Splat2DParams;

}  // namespace dsh
