// Differentiable layers, pipeline and Adam with the reference's class names,
// wiring and ownership rules, each method a thin call into the C-ABI.
//
//   Layer / TypedLayer       src/layers/layer.h:19-70
//   GaussActivLayer          src/layers/gauss_activ_layer.h:26-47
//   PerspProjectLayer        src/layers/persp_project_layer.h:25-38
//   TileRasterizeLayer       src/layers/tile_rasterize_layer.h:22-60
//   MSELossLayer             src/layers/mse_loss_layer.h:12-44
//   Pipeline                 src/pipelines/pipeline.h:5-31
//   Adam / AdamGroup         src/optimizers/adam.h:5-70
//   Gaussian3DAdamConfig, registerGaussian3DGroups  src/optimizers/gaussian3d_adam.h:11-55
//
// Ownership is the reference's: setInput/setGradOutput borrow, getOutput /
// getGradInput return owned buffers. Additions: every layer runs on a
// settable HIP stream; zeroGrad() only does work where backward()
// accumulates (the rasterizer) because every other backward overwrites its
// whole grad_input.
#pragma once

#include <array>
#include <cstring>
#include <vector>

#include "types.hpp"

namespace dsh {

class Layer {
public:
    virtual ~Layer() = default;
    virtual void forward() = 0;
    virtual void backward() = 0;
    virtual void zeroGrad() = 0;
    void setStream(hipStream_t s) { stream = s; }
    // DS_GRAD_EXACT (true derivative) or DS_GRAD_REFERENCE (the reference's expressions).
    void setGradMode(int m) { grad_mode = m; }

protected:
    hipStream_t stream = nullptr;
    int grad_mode = DS_GRAD_EXACT;
};

template <typename TIn, typename TOut, typename TGradIn, typename TGradOut>
class TypedLayer : public Layer {
public:
    void setInput(const TIn *in_) { input = in_; }
    TOut &getOutput() { return output; }
    void setGradOutput(const TGradOut *gout_) { grad_output = gout_; }
    TGradIn &getGradInput() { return grad_input; }

    virtual void allocate(int n) { output.allocate(n); }
    virtual void allocateGrad(int n) { grad_input.allocate(n); }
    void zeroGrad() override {}

protected:
    void require_wired(const char *what) const {
        if (!input) throw std::runtime_error(std::string("[diffsplat] ") + what + ": input not set");
    }
    const TIn *input = nullptr;
    TOut output;
    const TGradOut *grad_output = nullptr;
    TGradIn grad_input;
};

// ------------------------------------------------------------------ activation
class GaussActivLayer
    : public TypedLayer<Gaussian3DParams, Splat3DParams, Gaussian3DGrads, Splat3DGrads> {
public:
    void setSHDegree(int degree) { sh_degree = degree; }
    void setCameraPosition(float x, float y, float z) { cam = {x, y, z}; }
    void allocateGrad(int n) override { grad_input.allocate(n, sh_degree); }

    void forward() override {
        require_wired("GaussActivLayer::forward");
        output.count = input->count;
        ds_gauss3d in = input->view();
        ds_splat3d out = output.view();
        check(ds_activ_fwd(&in, &out, sh_degree, cam.data(), stream), "GaussActivLayer::forward");
    }
    void backward() override {
        ds_gauss3d in = input->view();
        ds_splat3d out = output.view();
        ds_splat3d go = grad_output->view();
        ds_gauss3d gi = grad_input.view();
        check(ds_activ_bwd(&in, &out, &go, &gi, sh_degree, cam.data(), grad_mode, stream),
              "GaussActivLayer::backward");
    }

private:
    int sh_degree = 0;
    std::array<float, 3> cam{0.f, 0.f, 0.f};
};

// ------------------------------------------------------------------ projection
class PerspProjectLayer
    : public TypedLayer<Splat3DParams, Splat2DParams, Splat3DGrads, Splat2DGrads> {
public:
    // PV = proj * view, both column-major.
    void setCamera(const float view[16], const float proj[16]) { ds_mat_mul(proj, view, pv); }
    const float *getPV() const { return pv; }

    void forward() override {
        require_wired("PerspProjectLayer::forward");
        output.count = input->count;
        ds_splat3d in = input->view();
        ds_splat2d out = output.view();
        check(ds_project_fwd(&in, &out, pv, stream), "PerspProjectLayer::forward");
    }
    void backward() override {
        ds_splat3d in = input->view();
        ds_splat2d go = grad_output->view();
        ds_splat3d gi = grad_input.view();
        check(ds_project_bwd(&in, &go, &gi, pv, grad_mode, stream),
              "PerspProjectLayer::backward");
    }

private:
    float pv[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

// Fused GaussActivLayer + PerspProjectLayer: params -> Splat2D in one kernel.
class GaussProjectLayer
    : public TypedLayer<Gaussian3DParams, Splat2DParams, Gaussian3DGrads, Splat2DGrads> {
public:
    void setSHDegree(int degree) { sh_degree = degree; }
    void setCameraPosition(float x, float y, float z) { cam = {x, y, z}; }
    void setCamera(const float view[16], const float proj[16]) { ds_mat_mul(proj, view, pv); }
    void allocateGrad(int n) override { grad_input.allocate(n, sh_degree); }

    void forward() override {
        require_wired("GaussProjectLayer::forward");
        output.count = input->count;
        ds_gauss3d in = input->view();
        ds_splat2d out = output.view();
        check(ds_activ_project_fwd(&in, &out, sh_degree, cam.data(), pv, stream),
              "GaussProjectLayer::forward");
    }
    void backward() override {
        ds_gauss3d in = input->view();
        ds_splat2d go = grad_output->view();
        ds_gauss3d gi = grad_input.view();
        check(ds_activ_project_bwd(&in, &go, &gi, sh_degree, cam.data(), pv, grad_mode, stream),
              "GaussProjectLayer::backward");
    }
    int shDegree() const { return sh_degree; }
    const std::array<float, 3> &cameraPosition() const { return cam; }
    const float *viewProj() const { return pv; }

private:
    int sh_degree = 0;
    std::array<float, 3> cam{0.f, 0.f, 0.f};
    float pv[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

// ------------------------------------------------------------------ rasterizer
class TileRasterizeLayer
    : public TypedLayer<Splat2DParams, DeviceBuffer<float>, Splat2DGrads, DeviceBuffer<float>> {
public:
    ~TileRasterizeLayer() override {
        if (ras) ds_raster_destroy(ras);
    }
    using TypedLayer::allocate;
    void allocate(int width, int height, int count, int64_t isect_cap = 0, int keep_keys = 0,
                  int grad_mode = DS_GRAD_EXACT) {
        if (ras) ds_raster_destroy(ras);
        ras = nullptr;
        ds_raster_config cfg{width, height, count, isect_cap, keep_keys, grad_mode};
        check(ds_raster_create(&cfg, &ras), "TileRasterizeLayer::allocate");
        w = width;
        h = height;
        output.allocate((size_t)width * height * 3);
    }
    // For forward-only use: frees the backward's per-pair gradient records.
    void releaseBackward() { check(ds_raster_release_backward(ras), "TileRasterizeLayer"); }
    void resize(int width, int height) {
        if (width == w && height == h) return;
        check(ds_raster_resize(ras, width, height), "TileRasterizeLayer::resize");
        w = width;
        h = height;
        output.allocate((size_t)width * height * 3);
    }
    // Fuses the MSE loss and its gradient into forward() (see ds_raster_fwd).
    void setLossTarget(const float *target, float *loss_dev, float *grad_dev) {
        loss_target = target;
        loss_out = loss_dev;
        loss_grad = grad_dev;
    }
    void zeroGrad() override {
        ds_splat2d gi = grad_input.view();
        check(ds_splat2d_zero(&gi, stream), "TileRasterizeLayer::zeroGrad");
    }
    // With the MSE fused into the forward (setLossTarget), the backward can
    // run inside the forward too (ds_raster_fwd_bwd): forward() then also
    // accumulates dL/dSplat2D and the next backward() does nothing. zeroGrad
    // must still come before forward().
    void setFuseBackward(bool on) { fuse_backward = on; }
    void forward() override {
        require_wired("TileRasterizeLayer::forward");
        ds_splat2d in = input->view();
        fused_last = fuse_backward && loss_target != nullptr;
        if (fused_last) {
            ds_splat2d gi = grad_input.view();
            check(ds_raster_fwd_bwd(ras, &in, output.ptr, loss_target, loss_out, loss_grad, &gi,
                                    stream),
                  "TileRasterizeLayer::forward");
        } else {
            check(ds_raster_fwd(ras, &in, output.ptr, loss_target, loss_out, loss_grad, stream),
                  "TileRasterizeLayer::forward");
        }
    }
    void backward() override {
        if (fused_last) {  // done by the forward
            fused_last = false;
            return;
        }
        ds_splat2d in = input->view();
        ds_splat2d gi = grad_input.view();
        check(ds_raster_bwd(ras, &in, grad_output->ptr, &gi, stream),
              "TileRasterizeLayer::backward");
    }
    uint32_t getVisibleCount() {
        uint32_t v = 0;
        check(ds_raster_counts(ras, stream, nullptr, &v, nullptr), "getVisibleCount");
        return v;
    }
    ds_raster *raster() const { return ras; }

private:
    ds_raster *ras = nullptr;
    int w = 0, h = 0;
    bool fuse_backward = false;
    bool fused_last = false;  // the last forward also ran the backward
    const float *loss_target = nullptr;
    float *loss_out = nullptr;
    float *loss_grad = nullptr;
};

// ------------------------------------------------------------------ loss
class MSELossLayer : public Layer {
public:
    void allocate(int width, int height) {
        num_pixels = (int64_t)width * height;
        d_grad_pixels.allocate((size_t)num_pixels * 3);
        d_loss.allocate(1);
        d_loss.zero(nullptr);
    }
    // Fused mode: the rasterizer's forward computes loss and gradient into
    // this layer's buffers and forward()/backward() have nothing left to do.
    void setFusedIntoRasterizer(bool on) { fused = on; }
    void forward() override {
        if (fused) return;
        check(ds_mse_fwd(d_in_pixels, d_target_pixels, num_pixels, d_loss.ptr, stream),
              "MSELossLayer::forward");
    }
    void backward() override {
        if (fused) return;
        check(ds_mse_bwd(d_in_pixels, d_target_pixels, num_pixels, d_grad_pixels.ptr, stream),
              "MSELossLayer::backward");
    }
    void zeroGrad() override {}
    void setInput(const float *pixels) { d_in_pixels = pixels; }
    void setTarget(const float *target) { d_target_pixels = target; }
    DeviceBuffer<float> &getGradInput() { return d_grad_pixels; }
    float *lossDevice() const { return d_loss.ptr; }
    float getLoss() const {
        float h = 0.f;
        if (stream) check(ds_stream_synchronize(stream), "MSELossLayer::getLoss");
        check(ds_memcpy_dtoh(&h, d_loss.ptr, sizeof(float)), "MSELossLayer::getLoss");
        return h;
    }

private:
    const float *d_in_pixels = nullptr;
    const float *d_target_pixels = nullptr;
    DeviceBuffer<float> d_loss;
    DeviceBuffer<float> d_grad_pixels;
    int64_t num_pixels = 0;
    bool fused = false;
};

// ------------------------------------------------------------------ pipeline
class Pipeline {
public:
    void add(Layer *layer) { layers.push_back(layer); }
    void clear() { layers.clear(); }
    void zeroGrad() {
        for (auto *l : layers) l->zeroGrad();
    }
    void forward() {
        for (auto *l : layers) l->forward();
    }
    void backward() {
        for (auto it = layers.rbegin(); it != layers.rend(); ++it) (*it)->backward();
    }

private:
    std::vector<Layer *> layers;
};

// ------------------------------------------------------------------ Adam
struct AdamParams {
    float beta1 = 0.9f;
    float beta2 = 0.999f;
    float epsilon = 1e-8f;
};

struct AdamGroup {
    float *param;
    const float *grad;
    int64_t n;
    float lr;
    float *m;  // views into the optimizer's moment slabs
    float *v;
    float decay;  // lr * decay^(t-1); 1: constant (the reference)
};

// One launch per step over every group; the step counter lives on the
// device so a captured graph replays correctly.
class Adam {
public:
    void init(const AdamParams &p = {}) { hp = p; }
    void addGroup(float *param, const float *grad, int64_t n, float lr, float decay = 1.f) {
        groups.push_back({param, grad, n, lr, nullptr, nullptr, decay});
        built = false;
    }
    void setLr(int idx, float lr) { groups[idx].lr = lr; }
    void setDecay(int idx, float decay) { groups[idx].decay = decay; }
    // Allocates the moment slabs and zeroes them and the step counter.
    void reset(hipStream_t s = nullptr) {
        int64_t total = 0;
        for (auto &g : groups) total += g.n;
        m.allocate((size_t)total);
        v.allocate((size_t)total);
        step_dev.allocate(1);
        int64_t off = 0;
        for (auto &g : groups) {
            g.m = m.ptr + off;
            g.v = v.ptr + off;
            off += g.n;
        }
        m.zero(s);
        v.zero(s);
        step_dev.zero(s);
        step_count = 0;
        built = true;
    }
    void step(hipStream_t s = nullptr) {
        if (!built) reset(s);
        std::vector<ds_adam_group> tab;
        tab.reserve(groups.size());
        for (auto &g : groups) tab.push_back({g.param, g.grad, g.m, g.v, g.n, g.lr, g.decay});
        ds_adam_hparams h{hp.beta1, hp.beta2, hp.epsilon};
        check(ds_adam_step_dev(tab.data(), (int)tab.size(), &h, step_dev.ptr, s), "Adam::step");
        ++step_count;
    }
    int getStepCount() const { return step_count; }
    const AdamParams &getParams() const { return hp; }
    int32_t *stepDevice() const { return step_dev.ptr; }
    // First and second moments of every group, in the order the groups were
    // added (one element per parameter).
    const DeviceBuffer<float> &moment1() const { return m; }
    const DeviceBuffer<float> &moment2() const { return v; }
    // The groups as the C-ABI takes them, for callers that launch the step
    // themselves (then countStep() keeps the host count).
    std::vector<ds_adam_group> table() {
        if (!built) throw std::runtime_error("[diffsplat] Adam::table before reset");
        std::vector<ds_adam_group> tab;
        tab.reserve(groups.size());
        for (auto &g : groups) tab.push_back({g.param, g.grad, g.m, g.v, g.n, g.lr, g.decay});
        return tab;
    }
    void countStep() { ++step_count; }
    void setStepCount(int t) { step_count = t; }

private:
    AdamParams hp;
    std::vector<AdamGroup> groups;
    DeviceBuffer<float> m, v;
    DeviceBuffer<int32_t> step_dev;
    int step_count = 0;
    bool built = false;
};

struct Gaussian3DAdamConfig {
    float lr_pos = 1.6e-2f;
    float lr_scale = 1.0e-3f;
    float lr_rot = 6.0e-4f;
    float lr_sh_dc = 6.0e-3f;
    float lr_sh_rest = 6.0e-3f;
    float lr_opacity = 6.0e-3f;
    float lr_pos_decay = 1.f;  // per-step position lr schedule (not in the reference)
    AdamParams adam = {};
};

inline void registerGaussian3DGroups(Adam &adam, Gaussian3DParams &p, Gaussian3DGrads &g,
                                     const Gaussian3DAdamConfig &cfg = {}) {
    const int64_t n = p.count;
    adam.addGroup(p.pos_x, g.pos_x, n, cfg.lr_pos, cfg.lr_pos_decay);
    adam.addGroup(p.pos_y, g.pos_y, n, cfg.lr_pos, cfg.lr_pos_decay);
    adam.addGroup(p.pos_z, g.pos_z, n, cfg.lr_pos, cfg.lr_pos_decay);
    adam.addGroup(p.scale_x, g.scale_x, n, cfg.lr_scale);
    adam.addGroup(p.scale_y, g.scale_y, n, cfg.lr_scale);
    adam.addGroup(p.scale_z, g.scale_z, n, cfg.lr_scale);
    adam.addGroup(p.rot_w, g.rot_w, n, cfg.lr_rot);
    adam.addGroup(p.rot_x, g.rot_x, n, cfg.lr_rot);
    adam.addGroup(p.rot_y, g.rot_y, n, cfg.lr_rot);
    adam.addGroup(p.rot_z, g.rot_z, n, cfg.lr_rot);
    adam.addGroup(p.sh_dc_r, g.sh_dc_r, n, cfg.lr_sh_dc);
    adam.addGroup(p.sh_dc_g, g.sh_dc_g, n, cfg.lr_sh_dc);
    adam.addGroup(p.sh_dc_b, g.sh_dc_b, n, cfg.lr_sh_dc);
    adam.addGroup(p.logit_opacity, g.logit_opacity, n, cfg.lr_opacity);
    if (p.sh_num_bands > 0) {
        const int64_t nsh = n * p.sh_num_bands;
        adam.addGroup(p.sh_rest_r, g.sh_rest_r, nsh, cfg.lr_sh_rest);
        adam.addGroup(p.sh_rest_g, g.sh_rest_g, nsh, cfg.lr_sh_rest);
        adam.addGroup(p.sh_rest_b, g.sh_rest_b, nsh, cfg.lr_sh_rest);
    }
}

}  // namespace dsh
