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
    // Adam moments m and v in the parameter slab's layout (either may be NULL).
    void downloadMoments(float *host_m, float *host_v) const;
    void initLayers();

    void step();          // zeroGrad, forward, backward, Adam
    void steps(int n);    // n steps (graph mode: replays of a multi-step graph)
    void prepare(int n);  // graph mode: capture what steps(n) replays, without running it
    void forwardOnly();   // forward (render + loss)
    void forwardBackward();  // zeroGrad, forward, backward (no Adam update)
    void setTargetDevice(const float *dev_rgb);  // device-to-device, on the fitter's stream
    void resetOptimizer();  // Adam moments and step counter to zero, parameters kept
    float *paramSlab() { return gaussian_params.slab.ptr; }
    const float *gradSlab();
    void setUseGraph(bool on);
    // Runs the fitter's work on stream s from now on (nullptr: its own
    // stream). Synchronises the old stream; a captured step is dropped and
    // captured again on the next step.
    void bindStream(hipStream_t s);
    // Times the compositing launch shapes on the current parameters (forward
    // and backward only, no Adam update) and keeps the fastest of each.
    void autotune(int reps, int *best_fwd, int *best_bwd, float *ms_fwd, float *ms_bwd);

    float getLoss();
    float *getOutput() { return ras_layer.getOutput().ptr; }
    int getIterCount() const { return iter_count; }
    int count() const { return opt.count; }
    int shDegree() const { return opt.sh_degree; }
    hipStream_t getStream() const { return stream; }
    ds_raster *raster() const { return ras_layer.raster(); }
    int width() const { return opt.width; }
    int height() const { return opt.height; }
    void views(ds_gauss3d *params, ds_splat2d *s2, ds_splat2d *gs2, float **grad_rgb);
    // Adaptive density control (kernels/densify.hip): per-step gradient
    // statistics, then one clone/split/prune with device-side compaction of
    // the parameters and Adam moments. The count changes; the step counter,
    // launch shapes and everything else are kept.
    void setDensifyStats(bool on);
    void downloadDensifyStats(float *grad_sum, uint32_t *visits);
    void densify(const ds_densify_config &cfg, ds_densify_result *res, uint8_t *host_actions);

private:
    void stepOnce();
    void accumulateDensifyStats();
    bool densify_stats = false;
    DeviceBuffer<float> dens_accum;
    DeviceBuffer<uint32_t> dens_visits;
    void dropGraph();
    FitterOptions opt;
    bool fused = true;
    hipStream_t own_stream = nullptr;  // created by init, destroyed with the fitter
    hipStream_t stream = nullptr;      // the stream work goes to (own_stream unless bound)
    Gaussian3DParams gaussian_params;
    Pipeline pipeline;
    GaussActivLayer atv_layer;
    PerspProjectLayer psp_layer;
    GaussProjectLayer gp_layer;
    TileRasterizeLayer ras_layer;
    MSELossLayer mse_layer;
    Adam optimizer;
    DeviceBuffer<float> d_target_pixels;
    int iter_count = 0;  // steps since the optimizer was last reset (graph replays too)
    bool use_graph = false;
    // captured steps: one step, and kGraphSteps steps (fewer graph launches)
    static constexpr int kGraphSteps = 16;
    hipGraphExec_t graph_exec = nullptr, graph_exec_k = nullptr;
    hipGraph_t graph = nullptr, graph_k = nullptr;
    void capture(int nsteps, hipGraph_t &g, hipGraphExec_t &ge);
    void launchCaptured(int nsteps, hipGraph_t &g, hipGraphExec_t &ge);
};

// Many independent fits stepped together (no counterpart in the reference):
// each instance replays its own captured step, the instances are spread over
// a few streams (largest measured step time first, each to the least-loaded
// stream) so that small fits overlap on the GPU's hardware queues. The
// fitters are borrowed and must outlive the batch.
class BatchFitter {
public:
    BatchFitter() = default;
    ~BatchFitter();
    BatchFitter(const BatchFitter &) = delete;
    BatchFitter &operator=(const BatchFitter &) = delete;

    void init(std::vector<ImageFitter *> fitters);
    // Spreads the instances over n_streams streams using the last measured
    // step times (equal weights before any measurement).
    void setStreams(int n_streams);
    int numStreams() const { return (int)streams.size(); }
    void step(int n);   // n steps of every instance (asynchronous)
    void synchronize();
    // Tunes each instance's compositing launch shapes, measures each
    // instance's step time, then times `reps` batch steps for 1, 2, 4 and 8
    // streams and keeps the fastest. Takes real optimisation steps.
    int autotune(int reps, float *ms_per_batch_step);
    int size() const { return (int)fit.size(); }
    ImageFitter *fitter(int i) { return fit[(size_t)i]; }

private:
    void releaseStreams();
    std::vector<ImageFitter *> fit;
    std::vector<hipStream_t> streams;
    std::vector<float> step_ms;  // measured step time per instance
};

// Many independent fits as one (no counterpart in the reference): every
// phase of a step (zero grads, activation + projection, tile assign, sorts,
// compositing with MSE, compositing backward, projection/activation
// backward, Adam) is one launch over all instances. The instances' splats
// are runs of one parameter slab, their cameras a device table, their tiles
// runs of one global tile space and their images runs of one pixel array
// (a multi-instance ds_raster). Per instance the step equals ImageFitter's
// with the same configuration. The step can be captured into one HIP graph.
struct MultiFitInstance {
    int width = 0, height = 0, count = 0;
    float view[16], proj[16];
    float cam_pos[3] = {0.f, 0.f, 0.f};
};

class MultiFitter {
public:
    MultiFitter() = default;
    ~MultiFitter();
    MultiFitter(const MultiFitter &) = delete;
    MultiFitter &operator=(const MultiFitter &) = delete;

    void init(const std::vector<MultiFitInstance> &instances, int sh_degree,
              const Gaussian3DAdamConfig &lr, int64_t isect_cap, int grad_mode);
    int size() const { return (int)inst.size(); }
    const MultiFitInstance &instance(int k) const { return inst[(size_t)k]; }
    void setTarget(int k, const float *host_rgb);
    // Instance k's slab in the ds_fitter_upload_params layout; resets the
    // optimizer of every instance (moments and the shared step counter).
    void uploadParams(int k, const float *host_slab);
    void downloadParams(int k, float *host_slab);
    void downloadGrads(int k, float *host_slab);
    void downloadImage(int k, float *host_rgb);
    void step();
    void forwardOnly();
    void forwardBackward();
    void resetOptimizer();
    void losses(float *host);  // one per instance, of the last forward (synchronises)
    void setUseGraph(bool on) {  // drops a captured step (see ImageFitter::setUseGraph)
        use_graph = on;
        dropGraph();
    }
    void autotune(int reps, int *best_fwd, int *best_bwd, float *ms_fwd, float *ms_bwd);
    int getIterCount() const { return iter_count; }
    hipStream_t getStream() const { return stream; }
    ds_raster *raster() const { return ras; }
    int64_t totalSplats() const { return n_total; }
    // Device views over all instances (instance k's splats are elements
    // [splat base, + count) of every field, its pixels [pixel base, + W H)).
    void views(ds_gauss3d *p, ds_splat2d *s2v, ds_splat2d *gs2v, float **grad) {
        if (p) *p = params.view();
        if (s2v) *s2v = s2.view();
        if (gs2v) *gs2v = gs2.view();
        if (grad) *grad = grad_rgb.ptr;
    }

private:
    void stepOnce();
    void dropGraph();
    void rowsCopy(int k, float *dev_slab, float *host, bool to_device);
    std::vector<MultiFitInstance> inst;
    std::vector<int64_t> splat_base, pix_base;
    int64_t n_total = 0, pix_total = 0;
    int sh_degree = 0;
    int grad_mode = DS_GRAD_EXACT;
    hipStream_t stream = nullptr;
    ds_raster *ras = nullptr;
    const uint16_t *splat_inst = nullptr;  // owned by ras
    Gaussian3DParams params;
    Gaussian3DGrads grads;
    Splat2DParams s2;
    Splat2DGrads gs2;
    DeviceBuffer<float> cams, target, image, grad_rgb, loss;
    Adam optimizer;
    int iter_count = 0;
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
    // The projected splats of the last frame (the compositing's input).
    ds_splat2d splat2dView() { return gp_layer.getOutput().view(); }
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
