// ImageFitter and SplatRenderer with the reference's step order and wiring.
//
//   ImageFitter     src/pipelines/image_fitter.{h,cpp}  (init 18-34, initLayers 63-99,
//                                                         step 103-112)
//   SplatRenderer   src/pipelines/splat_renderer.{h,cpp} (initLayers 50-75, render 85-91)
//
// Differences: the camera, target and initial parameters come in as arrays
// (image decoding, PLY and windowing are not part of this engine); the
// fitter can run fused (activation+projection in one kernel, MSE inside the
// compositing epilogue) and can capture its step into a HIP graph.
#pragma once

#include "layers.hpp"

namespace dsh {

struct FitterOptions {
    int width = 0, height = 0, count = 0, sh_degree = 0;
    float view[16], proj[16];
    float cam_pos[3] = {0.f, 0.f, 0.f};
    Gaussian3DAdamConfig lr;
    int64_t isect_cap = 0;
    bool fused = true;
    int grad_mode = DS_GRAD_EXACT;
};

class ImageFitter {
public:
    ImageFitter();
    ~ImageFitter();
    ImageFitter(const ImageFitter &) = delete;
    ImageFitter &operator=(const ImageFitter &) = delete;

    // Reference init: identity view and ortho(+-W/2, +-H/2, +-min(W,H)/8).
    static void referenceCamera(int w, int h, float view[16], float proj[16]);

    void init(const FitterOptions &opt);
    void setTarget(const float *host_rgb);
    void uploadParams(const float *host_slab);  // also resets Adam
    void downloadParams(float *host_slab) const;
    void downloadGrads(float *host_slab) const;
    void initLayers();

    void step();          // zeroGrad, forward, backward, Adam
    void forwardOnly();   // forward (render + loss)
    void setUseGraph(bool on);
    // Times the compositing launch shapes on the current parameters (forward
    // and backward only, no Adam update) and keeps the fastest of each.
    void autotune(int reps, int *best_fwd, int *best_bwd, float *ms_fwd, float *ms_bwd);

    float getLoss();
    float *getOutput() { return ras_layer.getOutput().ptr; }
    int getIterCount() const { return optimizer.getStepCount(); }
    hipStream_t getStream() const { return stream; }
    ds_raster *raster() const { return ras_layer.raster(); }
    int width() const { return opt.width; }
    int height() const { return opt.height; }
    void views(ds_gauss3d *params, ds_splat2d *s2, ds_splat2d *gs2, float **grad_rgb);

private:
    void stepOnce();
    FitterOptions opt;
    bool fused = true;
    hipStream_t stream = nullptr;
    Gaussian3DParams gaussian_params;
    Pipeline pipeline;
    GaussActivLayer atv_layer;
    PerspProjectLayer psp_layer;
    GaussProjectLayer gp_layer;
    TileRasterizeLayer ras_layer;
    MSELossLayer mse_layer;
    Adam optimizer;
    DeviceBuffer<float> d_target_pixels;
    bool use_graph = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
};

class SplatRenderer {
public:
    SplatRenderer();
    ~SplatRenderer();
    SplatRenderer(const SplatRenderer &) = delete;
    SplatRenderer &operator=(const SplatRenderer &) = delete;

    void init(int w, int h, int count, int sh_degree, int64_t isect_cap = 0);
    void uploadParams(const float *host_slab);
    void render(const float view[16], const float proj[16], const float cam_pos[3]);
    // Times the frame at the last camera for each compositing launch shape
    // (1, 2 or 4 pixels per thread), keeps the fastest and returns it.
    int autotune(int reps, float *ms_best);
    float *getOutput() { return ras_layer.getOutput().ptr; }
    uint32_t getVisibleCount() { return ras_layer.getVisibleCount(); }
    void setActiveSHDegree(int d) { gp_layer.setSHDegree(d); }
    int getMaxSHDegree() const { return sh_degree; }
    hipStream_t getStream() const { return stream; }
    ds_raster *raster() const { return ras_layer.raster(); }
    int width() const { return w; }
    int height() const { return h; }

private:
    int w = 0, h = 0, sh_degree = 0;
    hipStream_t stream = nullptr;
    Gaussian3DParams gaussian_params;
    Pipeline pipeline;
    GaussProjectLayer gp_layer;
    TileRasterizeLayer ras_layer;
};

}  // namespace dsh
