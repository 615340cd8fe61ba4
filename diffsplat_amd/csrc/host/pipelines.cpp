// ImageFitter / BatchFitter / SplatRenderer implementation (see pipelines.hpp).
#include "pipelines.hpp"

#include "../ds_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <chrono>

namespace dsh {

static void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("[diffsplat] ") + what + ": " +
                                 hipGetErrorString(e));
}

// Index into {1, 2, 4} pixels per thread of the launch shape to keep, given
// the times t[] of the three: the current shape stays unless another is at
// least 5% faster. Shapes that tie on the tuning frame can differ a lot
// later in a fit, so a near-tie is not worth a switch.
static int pick_with_margin(const float t[3], int current_ppt) {
    int best = current_ppt == 1 ? 0 : current_ppt == 2 ? 1 : 2;
    for (int i = 0; i < 3; ++i)
        if (t[i] < 0.95f * t[best]) best = i;
    return best;
}

// Times `timed()` (ms) with each compositing dispatch order of r and keeps
// the faster (the current order unless the other is at least 5% faster).
template <class F>
static void tune_tile_order(ds_raster *r, F &&timed) {
    int cur = DS_TILE_ORDER_HEAVY_FIRST;
    check(ds_raster_get_tile_order(r, &cur), "autotune");
    float t[2];
    for (int o = 0; o < 2; ++o) {
        check(ds_raster_set_tile_order(r, o), "autotune");
        t[o] = timed();
    }
    const int other = 1 - cur;
    check(ds_raster_set_tile_order(r, t[other] < 0.95f * t[cur] ? other : cur), "autotune");
}

// Times the depth-sort modes on `in` and keeps the fastest (see
// ds_raster_autotune_depth_sort).
static int tune_depth_sort(ds_raster *r, const ds_splat2d &in, hipStream_t s, int reps) {
    int mode = 0;
    check(ds_raster_autotune_depth_sort(r, &in, reps, s, &mode, nullptr), "autotune");
    return mode;
}

// With the forward and backward launch shapes chosen apart (bf, bb), times
// the whole compositing (timed_fwd_bwd, ms) for (bf, bb) and for every equal
// pair (1, 1), (2, 2), (4, 4), with which the forward and backward run in one
// fused kernel (whose speed the separate timings do not predict), and keeps
// the fastest (the separately chosen pair unless another is at least 5%
// faster). Returns the kept shapes in bf, bb.
template <class F>
static void tune_fused_shapes(ds_raster *r, int &bf, int &bb, F &&timed_fwd_bwd) {
    const int cand[4][2] = {{bf, bb}, {1, 1}, {2, 2}, {4, 4}};
    float t[4];
    for (int i = 0; i < 4; ++i) {
        check(ds_raster_set_tuning(r, cand[i][0], cand[i][1]), "autotune");
        t[i] = timed_fwd_bwd();
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (t[i] < 0.95f * t[best]) best = i;
    bf = cand[best][0];
    bb = cand[best][1];
    check(ds_raster_set_tuning(r, bf, bb), "autotune");
}

// ============================================================ ImageFitter

ImageFitter::ImageFitter() = default;

ImageFitter::~ImageFitter() {
    if (stream) (void)hipStreamSynchronize(stream);
    dropGraph();
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

void ImageFitter::dropGraph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graph_exec_k) (void)hipGraphExecDestroy(graph_exec_k);
    if (graph_k) (void)hipGraphDestroy(graph_k);
    graph_exec = graph_exec_k = nullptr;
    graph = graph_k = nullptr;
}

void ImageFitter::bindStream(hipStream_t s) {
    if (stream) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    stream = s ? s : own_stream;
    Layer *all[5] = {&atv_layer, &psp_layer, &gp_layer, &ras_layer, &mse_layer};
    for (Layer *l : all) l->setStream(stream);
    dropGraph();  // the captured step launches on the old stream
}

void ImageFitter::referenceCamera(int w, int h, float view[16], float proj[16]) {
    const float hw = w * 0.5f, hh = h * 0.5f;
    const float zr = std::min(hw, hh) * 0.25f;
    for (int k = 0; k < 16; ++k) view[k] = (k % 5 == 0) ? 1.f : 0.f;
    ds_mat_ortho(-hw, hw, -hh, hh, -zr, zr, proj);
}

void ImageFitter::init(const FitterOptions &o) {
    if (o.width <= 0 || o.height <= 0 || o.count < 0 || o.sh_degree < 0 || o.sh_degree > 3)
        throw std::runtime_error("[diffsplat] ImageFitter::init: bad configuration");
    opt = o;
    fused = o.fused;
    if (!own_stream) hip_check(hipStreamCreate(&own_stream), "hipStreamCreate");
    if (!stream) stream = own_stream;
    gaussian_params.allocate(o.count, o.sh_degree);
    d_target_pixels.allocate((size_t)o.width * o.height * 3);
    d_target_pixels.zero(stream);
}

void ImageFitter::setTarget(const float *host_rgb) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    d_target_pixels.upload(host_rgb, (size_t)opt.width * opt.height * 3);
}

void ImageFitter::uploadParams(const float *host_slab) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    gaussian_params.slab.upload(host_slab, gaussian_params.slab.size);
    optimizer.reset(stream);
    iter_count = 0;
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void ImageFitter::downloadParams(float *host_slab) const {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    gaussian_params.slab.download(host_slab, gaussian_params.slab.size);
}

void ImageFitter::downloadGrads(float *host_slab) const {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const Gaussian3DGrads &g = fused ? const_cast<GaussProjectLayer &>(gp_layer).getGradInput()
                                     : const_cast<GaussActivLayer &>(atv_layer).getGradInput();
    g.slab.download(host_slab, g.slab.size);
}

void ImageFitter::downloadMoments(float *host_m, float *host_v) const {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // registerGaussian3DGroups adds the 13 leading fields, then the opacity,
    // then the SH rest bands; the slab keeps the opacity last.
    const size_t n = (size_t)opt.count;
    const size_t rest = 3 * n * (size_t)ds_sh_bands(opt.sh_degree);
    std::vector<float> tmp(optimizer.moment1().size);
    for (int which = 0; which < 2; ++which) {
        float *dst = which ? host_v : host_m;
        if (!dst) continue;
        (which ? optimizer.moment2() : optimizer.moment1()).download(tmp.data(), tmp.size());
        std::copy(tmp.begin(), tmp.begin() + 13 * n, dst);
        std::copy(tmp.begin() + 14 * n, tmp.begin() + 14 * n + rest, dst + 13 * n);
        std::copy(tmp.begin() + 13 * n, tmp.begin() + 14 * n, dst + 13 * n + rest);
    }
}

void ImageFitter::initLayers() {
    const int n = opt.count;
    pipeline.clear();
    Layer *all[5] = {&atv_layer, &psp_layer, &gp_layer, &ras_layer, &mse_layer};
    for (Layer *l : all) {
        l->setStream(stream);
        l->setGradMode(opt.grad_mode);
    }

    ras_layer.allocate(opt.width, opt.height, n, opt.isect_cap, 0, opt.grad_mode);
    ras_layer.allocateGrad(n);
    mse_layer.allocate(opt.width, opt.height);
    mse_layer.setTarget(d_target_pixels.ptr);
    mse_layer.setInput(ras_layer.getOutput().ptr);
    ras_layer.setGradOutput(&mse_layer.getGradInput());

    Gaussian3DGrads *param_grads = nullptr;
    if (fused) {
        gp_layer.setSHDegree(opt.sh_degree);
        gp_layer.setCamera(opt.view, opt.proj);
        gp_layer.setCameraPosition(opt.cam_pos[0], opt.cam_pos[1], opt.cam_pos[2]);
        gp_layer.allocate(n);
        gp_layer.allocateGrad(n);
        gp_layer.setInput(&gaussian_params);
        ras_layer.setInput(&gp_layer.getOutput());
        gp_layer.setGradOutput(&ras_layer.getGradInput());
        mse_layer.setFusedIntoRasterizer(true);
        ras_layer.setLossTarget(d_target_pixels.ptr, mse_layer.lossDevice(),
                                mse_layer.getGradInput().ptr);
        ras_layer.setFuseBackward(true);
        pipeline.add(&gp_layer);
        pipeline.add(&ras_layer);
        pipeline.add(&mse_layer);
        param_grads = &gp_layer.getGradInput();
    } else {
        atv_layer.setSHDegree(opt.sh_degree);
        atv_layer.setCameraPosition(opt.cam_pos[0], opt.cam_pos[1], opt.cam_pos[2]);
        psp_layer.setCamera(opt.view, opt.proj);
        atv_layer.allocate(n);
        psp_layer.allocate(n);
        atv_layer.allocateGrad(n);
        psp_layer.allocateGrad(n);
        atv_layer.setInput(&gaussian_params);
        psp_layer.setInput(&atv_layer.getOutput());
        ras_layer.setInput(&psp_layer.getOutput());
        psp_layer.setGradOutput(&ras_layer.getGradInput());
        atv_layer.setGradOutput(&psp_layer.getGradInput());
        mse_layer.setFusedIntoRasterizer(false);
        ras_layer.setLossTarget(nullptr, nullptr, nullptr);
        ras_layer.setFuseBackward(false);
        pipeline.add(&atv_layer);
        pipeline.add(&psp_layer);
        pipeline.add(&ras_layer);
        pipeline.add(&mse_layer);
        param_grads = &atv_layer.getGradInput();
    }
    param_grads->zero(stream);
    optimizer = Adam();
    optimizer.init(opt.lr.adam);
    registerGaussian3DGroups(optimizer, gaussian_params, *param_grads, opt.lr);
    optimizer.reset(stream);
    iter_count = 0;
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// The optimiser's groups as the fused projection backward's fields (see
// ds::AdamFields): false unless every field of the slab (and nothing else) is
// a group with that field's parameters and gradients.
static bool adam_fields_of(const std::vector<ds_adam_group> &tab, const ds_gauss3d &p,
                           const ds_gauss3d &g, int deg, const AdamParams &hp,
                           int32_t *step, ds::AdamFields &af) {
    const float *pf[ds::kAdamFields] = {
        p.pos_x,   p.pos_y,   p.pos_z,   p.scale_x,   p.scale_y,   p.scale_z,
        p.rot_w,   p.rot_x,   p.rot_y,   p.rot_z,     p.sh_dc_r,   p.sh_dc_g,
        p.sh_dc_b, p.logit_opacity, p.sh_rest_r, p.sh_rest_g, p.sh_rest_b};
    const float *gf[ds::kAdamFields] = {
        g.pos_x,   g.pos_y,   g.pos_z,   g.scale_x,   g.scale_y,   g.scale_z,
        g.rot_w,   g.rot_x,   g.rot_y,   g.rot_z,     g.sh_dc_r,   g.sh_dc_g,
        g.sh_dc_b, g.logit_opacity, g.sh_rest_r, g.sh_rest_g, g.sh_rest_b};
    const int nfields = deg > 0 ? ds::kAdamFields : 14;
    if ((int)tab.size() != nfields) return false;
    af = ds::AdamFields{};
    for (int f = 0; f < nfields; ++f) {
        const int64_t n = (int64_t)p.count * (f < 14 ? 1 : ds_sh_bands(deg));
        const ds_adam_group *hit = nullptr;
        for (const auto &k : tab)
            if (k.param == pf[f] && k.grad == gf[f] && k.n == n) hit = &k;
        if (!hit) return false;
        af.p[f] = hit->param;
        af.m[f] = hit->m;
        af.v[f] = hit->v;
        af.lr[f] = hit->lr;
        af.decay[f] = hit->decay > 0.f ? hit->decay : 1.f;
    }
    af.beta1 = hp.beta1;
    af.beta2 = hp.beta2;
    af.eps = hp.epsilon;
    af.step = step;
    return true;
}

void ImageFitter::stepOnce() {
    if (fused) {
        // The reference's step (zeroGrad, forward, backward, Adam) with the
        // zeroing of the Splat2D gradients and the Adam step counter folded
        // into the rasterizer's first kernel, the compositing forward, MSE
        // and compositing backward in one kernel, and Adam in the
        // projection backward; the rasterizer's first-kernel work (scalars,
        // counters, that prelude) runs in the activation/projection kernel:
        // 13 launches.
        ds_splat2d gi = ras_layer.getGradInput().view();
        const int n = gaussian_params.count;
        bool started = false;
        if (n > 0) {
            ds::ForwardStart fs;
            check(ds::raster_step_start(ras_layer.raster(), n, mse_layer.lossDevice(), gi,
                                        optimizer.stepDevice(), &fs),
                  "ImageFitter::step");
            gp_layer.getOutput().count = n;
            const auto &c = gp_layer.cameraPosition();
            ds::PV pv;
            std::memcpy(pv.m, gp_layer.viewProj(), sizeof(pv.m));
            check(ds::launch_activ_project_fwd_start(gaussian_params.view(),
                                                     gp_layer.getOutput().view(),
                                                     gp_layer.shDegree(),
                                                     ds::Cam3{c[0], c[1], c[2]}, pv, fs, stream),
                  "ImageFitter::step");
            started = true;
        } else {
            gp_layer.forward();
        }
        ds_splat2d in = gp_layer.getOutput().view();
        check(ds::raster_step_fwd_bwd(ras_layer.raster(), in, ras_layer.getOutput().ptr,
                                      d_target_pixels.ptr, mse_layer.lossDevice(),
                                      mse_layer.getGradInput().ptr, gi, optimizer.stepDevice(),
                                      stream, started),
              "ImageFitter::step");
        std::vector<ds_adam_group> tab = optimizer.table();
        const AdamParams &a = optimizer.getParams();
        const ds_gauss3d params = gaussian_params.view();
        const ds_gauss3d pgrads = gp_layer.getGradInput().view();
        ds::AdamFields af;
        if (adam_fields_of(tab, params, pgrads, gp_layer.shDegree(), a, optimizer.stepDevice(),
                           af)) {
            const auto &c = gp_layer.cameraPosition();
            ds::PV pv;
            std::memcpy(pv.m, gp_layer.viewProj(), sizeof(pv.m));
            check(ds::launch_activ_project_bwd_adam(params, gi, pgrads, gp_layer.shDegree(),
                                                    ds::Cam3{c[0], c[1], c[2]}, pv,
                                                    opt.grad_mode, af, stream),
                  "ImageFitter::step");
        } else {
            gp_layer.backward();
            ds_adam_hparams h{a.beta1, a.beta2, a.epsilon};
            check(ds::launch_adam(tab.data(), (int)tab.size(), h, 0, optimizer.stepDevice(),
                                  stream, false),
                  "ImageFitter::step");
        }
        optimizer.countStep();
        accumulateDensifyStats();
        return;
    }
    pipeline.zeroGrad();
    pipeline.forward();
    pipeline.backward();
    optimizer.step(stream);
    accumulateDensifyStats();
}

void ImageFitter::accumulateDensifyStats() {
    if (!densify_stats || gaussian_params.count <= 0) return;
    const ds_splat2d gi = ras_layer.getGradInput().view();
    check(ds::launch_densify_accum(gi.pos_x, gi.pos_y, ds::raster_splat_tiles(ras_layer.raster()),
                                   dens_accum.ptr, dens_visits.ptr, gaussian_params.count, stream),
          "ImageFitter::step");
}

void ImageFitter::setDensifyStats(bool on) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    densify_stats = on;
    const size_t n = (size_t)gaussian_params.count;
    dens_accum.allocate(n);
    dens_visits.allocate(n);
    dens_accum.zero(stream);
    dens_visits.zero(stream);
    dropGraph();  // a captured step has or lacks the accumulation
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void ImageFitter::downloadDensifyStats(float *grad_sum, uint32_t *visits) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const size_t n = (size_t)gaussian_params.count;
    if (dens_accum.size != n)
        throw std::runtime_error("[diffsplat] densify statistics are not enabled");
    if (grad_sum) dens_accum.download(grad_sum, n);
    if (visits) dens_visits.download(visits, n);
}

void ImageFitter::densify(const ds_densify_config &cfg, ds_densify_result *res,
                          uint8_t *host_actions) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const int n = gaussian_params.count;
    const int nf = gaussian_params.nfields();
    if (dens_accum.size != (size_t)n)
        throw std::runtime_error("[diffsplat] densify needs ds_fitter_set_densify_stats(f, 1) "
                                 "and some steps first");
    if (cfg.percent_dense < 0.f || cfg.min_opacity < 0.f || cfg.min_opacity >= 1.f ||
        cfg.split_scale_div <= 0.f || cfg.max_scale_frac < 0.f)
        throw std::runtime_error("[diffsplat] densify: bad configuration");
    if ((int64_t)2 * n >= ((int64_t)1 << 30))
        throw std::runtime_error("[diffsplat] densify: too many splats");
    // scene extent: the largest distance of a centre from their mean
    float extent = cfg.scene_extent;
    if (!(extent > 0.f) && n > 0) {
        std::vector<float> pos((size_t)3 * n);
        hip_check(hipMemcpy(pos.data(), gaussian_params.slab.ptr, sizeof(float) * 3 * n,
                            hipMemcpyDeviceToHost),
                  "hipMemcpy");
        double c[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n; ++i) c[d] += pos[(size_t)d * n + i];
        for (int d = 0; d < 3; ++d) c[d] /= n;
        double r2 = 0;
        for (int i = 0; i < n; ++i) {
            double q = 0;
            for (int d = 0; d < 3; ++d) {
                const double x = pos[(size_t)d * n + i] - c[d];
                q += x * x;
            }
            r2 = std::max(r2, q);
        }
        extent = (float)std::sqrt(r2);
    }
    ds::DensifyParams p{};
    p.grad_threshold = cfg.grad_threshold;
    p.clone_max_log_scale = cfg.percent_dense > 0.f && extent > 0.f
                                ? (float)std::log((double)cfg.percent_dense * extent)
                                : -INFINITY;
    p.min_logit = cfg.min_opacity > 0.f
                      ? (float)std::log((double)cfg.min_opacity / (1.0 - cfg.min_opacity))
                      : -INFINITY;
    p.prune_big = cfg.max_scale_frac > 0.f && extent > 0.f;
    p.prune_max_log_scale = p.prune_big ? (float)std::log((double)cfg.max_scale_frac * extent)
                                        : INFINITY;
    p.split_log_div = (float)std::log((double)cfg.split_scale_div);
    p.seed = cfg.seed;
    // classify, scan and scatter on the device
    DeviceBuffer<uint8_t> action;
    DeviceBuffer<uint32_t> keep, added, keep_off, add_off, words, part;
    DeviceBuffer<float> np, nm, nv;
    action.allocate((size_t)std::max(n, 1));
    for (auto *b : {&keep, &added, &keep_off, &add_off}) b->allocate((size_t)std::max(n, 1));
    words.allocate(4);
    part.allocate(8192);
    const size_t cap = (size_t)2 * std::max(n, 1) * nf;
    np.allocate(cap);
    nm.allocate(cap);
    nv.allocate(cap);
    words.zero(stream);
    const ds::DensifyScratch scr{action.ptr, keep.ptr,  added.ptr,     keep_off.ptr,
                                 add_off.ptr, words.ptr, words.ptr + 2, part.ptr};
    check(ds::launch_densify(gaussian_params.view(), optimizer.moment1().ptr,
                             optimizer.moment2().ptr, nf, dens_accum.ptr, dens_visits.ptr, p, scr,
                             np.ptr, nm.ptr, nv.ptr, stream),
          "ImageFitter::densify");
    uint32_t tot[2] = {0, 0};
    std::vector<uint8_t> act((size_t)n);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (n > 0) {
        words.download(tot, 2);
        action.download(act.data(), (size_t)n);
    }
    const int n_out = (int)(tot[0] + tot[1]);
    ds_densify_result r{n, n_out, 0, 0, 0};
    for (uint8_t a : act) {
        r.pruned += a == DS_DENSIFY_PRUNE;
        r.cloned += a == DS_DENSIFY_CLONE;
        r.split += a == DS_DENSIFY_SPLIT;
    }
    if (host_actions && n > 0) std::memcpy(host_actions, act.data(), (size_t)n);
    // rebuild the fitter at the new count, keeping step, shapes and order
    int fwd_ppt = 1, bwd_ppt = 1, order = DS_TILE_ORDER_HEAVY_FIRST;
    check(ds_raster_get_tuning(ras_layer.raster(), &fwd_ppt, &bwd_ppt), "densify");
    check(ds_raster_get_tile_order(ras_layer.raster(), &order), "densify");
    int32_t step_now = 0;
    hip_check(hipMemcpy(&step_now, optimizer.stepDevice(), sizeof(step_now),
                        hipMemcpyDeviceToHost),
              "hipMemcpy");
    const int iters = iter_count;
    dropGraph();
    opt.count = n_out;
    gaussian_params.allocate(n_out, opt.sh_degree);
    const size_t nfl = (size_t)nf * n_out;
    if (nfl)
        hip_check(hipMemcpyAsync(gaussian_params.slab.ptr, np.ptr, sizeof(float) * nfl,
                                 hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync");
    initLayers();  // new buffers, Adam groups at the new count (moments zero)
    check(ds_raster_set_tuning(ras_layer.raster(), fwd_ppt, bwd_ppt), "densify");
    check(ds_raster_set_tile_order(ras_layer.raster(), order), "densify");
    if (nfl) {
        hip_check(hipMemcpyAsync(optimizer.moment1().ptr, nm.ptr, sizeof(float) * nfl,
                                 hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(optimizer.moment2().ptr, nv.ptr, sizeof(float) * nfl,
                                 hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync");
    }
    hip_check(hipMemcpyAsync(optimizer.stepDevice(), &step_now, sizeof(step_now),
                             hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync");
    optimizer.setStepCount(step_now);
    iter_count = iters;
    dens_accum.allocate((size_t)n_out);
    dens_visits.allocate((size_t)n_out);
    dens_accum.zero(stream);
    dens_visits.zero(stream);
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (res) *res = r;
}

// A captured step keeps the launch shapes and dispatch order it was captured
// with, so (re)setting the mode drops it; the next step captures again.
void ImageFitter::setUseGraph(bool on) {
    use_graph = on;
    dropGraph();
}

// Captures nsteps steps into a graph (recorded, not run) unless captured.
// The optimizer's host step count is left as it was: the captured steps run
// when the graph is launched.
void ImageFitter::capture(int nsteps, hipGraph_t &g, hipGraphExec_t &ge) {
    if (ge) return;
    const int host_steps = optimizer.getStepCount();
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
    try {
        for (int k = 0; k < nsteps; ++k) stepOnce();
    } catch (...) {
        hipGraph_t bad = nullptr;
        (void)hipStreamEndCapture(stream, &bad);
        if (bad) (void)hipGraphDestroy(bad);
        optimizer.setStepCount(host_steps);
        throw;
    }
    optimizer.setStepCount(host_steps);
    hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    // its device-side setup now, not at its first launch
    hip_check(hipGraphUpload(ge, stream), "hipGraphUpload");
}

// Replays nsteps captured steps (captured on first use).
void ImageFitter::launchCaptured(int nsteps, hipGraph_t &g, hipGraphExec_t &ge) {
    capture(nsteps, g, ge);
    hip_check(hipGraphLaunch(ge, stream), "hipGraphLaunch");
    for (int k = 0; k < nsteps; ++k) optimizer.countStep();
}

void ImageFitter::step() { steps(1); }

void ImageFitter::steps(int n) {
    if (n <= 0) return;
    iter_count += n;
    if (!use_graph) {
        for (int k = 0; k < n; ++k) stepOnce();
        return;
    }
    // Each graph launch costs the GPU a few microseconds between graphs, so
    // long runs replay a graph of kGraphSteps steps.
    for (; n >= kGraphSteps; n -= kGraphSteps) launchCaptured(kGraphSteps, graph_k, graph_exec_k);
    for (; n > 0; --n) launchCaptured(1, graph, graph_exec);
}

// Captures (without running) the graphs that steps(n) replays, so that a
// later steps(n) pays no capture or instantiation.
void ImageFitter::prepare(int n) {
    if (!use_graph || n <= 0) return;
    if (n >= kGraphSteps) capture(kGraphSteps, graph_k, graph_exec_k);
    if (n % kGraphSteps) capture(1, graph, graph_exec);
}

void ImageFitter::forwardOnly() {
    // a forward alone must not run the fused backward
    ras_layer.setFuseBackward(false);
    pipeline.forward();
    ras_layer.setFuseBackward(fused);
}

void ImageFitter::forwardBackward() {
    pipeline.zeroGrad();
    pipeline.forward();
    pipeline.backward();
}

void ImageFitter::setTargetDevice(const float *dev_rgb) {
    hip_check(hipMemcpyAsync(d_target_pixels.ptr, dev_rgb,
                             sizeof(float) * (size_t)opt.width * opt.height * 3,
                             hipMemcpyDeviceToDevice, stream),
              "hipMemcpyAsync");
}

void ImageFitter::resetOptimizer() {
    optimizer.reset(stream);
    iter_count = 0;
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

const float *ImageFitter::gradSlab() {
    return fused ? gp_layer.getGradInput().slab.ptr : atv_layer.getGradInput().slab.ptr;
}

void ImageFitter::autotune(int reps, int *best_fwd, int *best_bwd, float *ms_fwd,
                           float *ms_bwd) {
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    auto timed = [&](auto &&body) {
        body();  // warm-up
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        for (int k = 0; k < reps; ++k) body();
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / reps;
    };
    ds_raster *r = ras_layer.raster();
    int fwd0 = 1, bwd0 = 1;
    check(ds_raster_get_tuning(r, &fwd0, &bwd0), "autotune");
    const int opts[3] = {1, 2, 4};
    float tfw[3], tbw[3];
    ras_layer.setFuseBackward(false);  // forward and backward timed apart
    for (int i = 0; i < 3; ++i) {
        check(ds_raster_set_tuning(r, opts[i], bwd0), "autotune");
        tfw[i] = timed([&] { pipeline.forward(); });
    }
    int bf = opts[pick_with_margin(tfw, fwd0)];
    check(ds_raster_set_tuning(r, bf, bwd0), "autotune");
    pipeline.forward();
    for (int i = 0; i < 3; ++i) {
        check(ds_raster_set_tuning(r, bf, opts[i]), "autotune");
        tbw[i] = timed([&] { ras_layer.backward(); });
    }
    const int ib = pick_with_margin(tbw, bwd0);
    int bb = opts[ib];
    const float tf = tfw[bf == 1 ? 0 : bf == 2 ? 1 : 2], tb = tbw[ib];
    check(ds_raster_set_tuning(r, bf, bb), "autotune");
    // the step's compositing: fused when the shapes are equal (fused mode)
    ras_layer.setFuseBackward(fused);
    auto fwd_bwd = [&] {
        return timed([&] {
            pipeline.forward();
            ras_layer.backward();
        });
    };
    if (fused) tune_fused_shapes(r, bf, bb, fwd_bwd);
    tune_tile_order(r, fwd_bwd);
    pipeline.forward();
    tune_depth_sort(r, fused ? gp_layer.getOutput().view() : psp_layer.getOutput().view(), stream,
                    reps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dropGraph();  // the captured step (if any) used the old launch shapes
    if (best_fwd) *best_fwd = bf;
    if (best_bwd) *best_bwd = bb;
    if (ms_fwd) *ms_fwd = tf;
    if (ms_bwd) *ms_bwd = tb;
}

void ImageFitter::views(ds_gauss3d *params, ds_splat2d *s2, ds_splat2d *gs2, float **grad_rgb) {
    if (params) *params = gaussian_params.view();
    if (s2) *s2 = fused ? gp_layer.getOutput().view() : psp_layer.getOutput().view();
    if (gs2) *gs2 = ras_layer.getGradInput().view();
    if (grad_rgb) *grad_rgb = mse_layer.getGradInput().ptr;
}

float ImageFitter::getLoss() { return mse_layer.getLoss(); }

// ============================================================ BatchFitter

BatchFitter::~BatchFitter() { releaseStreams(); }

void BatchFitter::releaseStreams() {
    for (ImageFitter *f : fit) f->bindStream(nullptr);  // synchronises the shared streams
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    streams.clear();
}

void BatchFitter::init(std::vector<ImageFitter *> fitters) {
    releaseStreams();
    fit = std::move(fitters);
    step_ms.assign(fit.size(), 1.f);
    for (ImageFitter *f : fit) f->setUseGraph(true);
    setStreams(1);
}

void BatchFitter::setStreams(int n_streams) {
    if (n_streams < 1) n_streams = 1;
    if (n_streams > (int)fit.size()) n_streams = std::max(1, (int)fit.size());
    releaseStreams();
    streams.resize((size_t)n_streams);
    for (hipStream_t &s : streams) hip_check(hipStreamCreate(&s), "hipStreamCreate");
    // longest step first, each to the stream with the least work so far
    std::vector<int> order(fit.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return step_ms[(size_t)a] > step_ms[(size_t)b]; });
    std::vector<float> load((size_t)n_streams, 0.f);
    for (int i : order) {
        const size_t s = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        load[s] += step_ms[(size_t)i];
        fit[(size_t)i]->bindStream(streams[s]);
    }
}

void BatchFitter::step(int n) {
    for (int k = 0; k < n; ++k)
        for (ImageFitter *f : fit) f->step();
}

void BatchFitter::synchronize() {
    for (hipStream_t s : streams) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int BatchFitter::autotune(int reps, float *ms_per_batch_step) {
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    // each instance alone on its own stream: launch shapes, then its step time
    for (size_t i = 0; i < fit.size(); ++i) {
        ImageFitter *f = fit[i];
        f->bindStream(nullptr);
        f->forwardOnly();
        f->autotune(reps, nullptr, nullptr, nullptr, nullptr);
        f->step();  // capture
        hip_check(hipEventRecord(e0, f->getStream()), "hipEventRecord");
        for (int k = 0; k < reps; ++k) f->step();
        hip_check(hipEventRecord(e1, f->getStream()), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        step_ms[i] = ms / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // stream count: wall time of `reps` batch steps
    int best = 1;
    float tbest = 1e30f;
    for (int ns : {1, 2, 4, 8}) {
        if (ns > 1 && ns > (int)fit.size()) break;
        setStreams(ns);
        step(1);  // captures on the new streams
        synchronize();
        const auto t0 = std::chrono::steady_clock::now();
        step(reps);
        synchronize();
        const float ms =
            std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count() /
            reps;
        if (ms < 0.95f * tbest) {
            tbest = ms;
            best = ns;
        }
    }
    setStreams(best);
    if (ms_per_batch_step) *ms_per_batch_step = tbest;
    return best;
}

// ============================================================ MultiFitter

MultiFitter::~MultiFitter() {
    if (stream) (void)hipStreamSynchronize(stream);
    dropGraph();
    if (ras) ds_raster_destroy(ras);
    if (stream) (void)hipStreamDestroy(stream);
}

void MultiFitter::dropGraph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

void MultiFitter::init(const std::vector<MultiFitInstance> &instances, int deg,
                       const Gaussian3DAdamConfig &lr, int64_t isect_cap, int mode) {
    if (instances.empty() || deg < 0 || deg > 3)
        throw std::runtime_error("[diffsplat] MultiFitter::init: bad configuration");
    inst = instances;
    sh_degree = deg;
    grad_mode = mode;
    splat_base.clear();
    pix_base.clear();
    n_total = pix_total = 0;
    std::vector<ds_raster_instance> ri;
    std::vector<float> cam_tab;
    for (const MultiFitInstance &q : inst) {
        if (q.width <= 0 || q.height <= 0 || q.count < 0)
            throw std::runtime_error("[diffsplat] MultiFitter::init: bad instance");
        splat_base.push_back(n_total);
        pix_base.push_back(pix_total);
        n_total += q.count;
        pix_total += (int64_t)q.width * q.height;
        ri.push_back({q.width, q.height, q.count});
        float pv[16];
        ds_mat_mul(q.proj, q.view, pv);
        cam_tab.insert(cam_tab.end(), pv, pv + 16);
        cam_tab.insert(cam_tab.end(), q.cam_pos, q.cam_pos + 3);
        cam_tab.push_back(0.f);
    }
    if (n_total > (1 << 30) - 1)
        throw std::runtime_error("[diffsplat] MultiFitter::init: too many splats");
    if (!stream) hip_check(hipStreamCreate(&stream), "hipStreamCreate");
    dropGraph();
    if (ras) ds_raster_destroy(ras);
    ras = nullptr;
    check(ds_raster_create_multi(ri.data(), (int)ri.size(), isect_cap, 0, grad_mode, &ras),
          "MultiFitter::init");
    check(ds_raster_splat_instances(ras, &splat_inst), "MultiFitter::init");
    const int n = (int)n_total;
    params.allocate(n, deg);
    grads.allocate(n, deg);
    s2.allocate(n);
    gs2.allocate(n);
    params.zero(stream);
    grads.zero(stream);
    cams.allocate(cam_tab.size());
    cams.upload(cam_tab.data(), cam_tab.size());
    target.allocate((size_t)pix_total * 3);
    image.allocate((size_t)pix_total * 3);
    grad_rgb.allocate((size_t)pix_total * 3);
    loss.allocate(inst.size());
    target.zero(stream);
    loss.zero(stream);
    optimizer = Adam();
    optimizer.init(lr.adam);
    registerGaussian3DGroups(optimizer, params, grads, lr);
    optimizer.reset(stream);
    iter_count = 0;
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// Instance k's rows (14 + 3 b of count floats each, the ds_fitter slab
// layout) to or from the combined slab, whose rows hold n_total floats.
void MultiFitter::rowsCopy(int k, float *dev_slab, float *host, bool to_device) {
    if (k < 0 || k >= size()) throw std::runtime_error("[diffsplat] MultiFitter: bad instance");
    const size_t nk = (size_t)inst[(size_t)k].count;
    if (nk == 0) return;
    const size_t rows = (size_t)ds_slab_floats(1, sh_degree);
    float *dev = dev_slab + splat_base[(size_t)k];
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (to_device)
        hip_check(hipMemcpy2D(dev, (size_t)n_total * 4, host, nk * 4, nk * 4, rows,
                              hipMemcpyHostToDevice),
                  "hipMemcpy2D");
    else
        hip_check(hipMemcpy2D(host, nk * 4, dev, (size_t)n_total * 4, nk * 4, rows,
                              hipMemcpyDeviceToHost),
                  "hipMemcpy2D");
}

void MultiFitter::setTarget(int k, const float *host_rgb) {
    if (k < 0 || k >= size()) throw std::runtime_error("[diffsplat] MultiFitter: bad instance");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const MultiFitInstance &q = inst[(size_t)k];
    hip_check(hipMemcpy(target.ptr + 3 * pix_base[(size_t)k], host_rgb,
                        sizeof(float) * 3 * (size_t)q.width * q.height, hipMemcpyHostToDevice),
              "hipMemcpy");
}

void MultiFitter::uploadParams(int k, const float *host_slab) {
    rowsCopy(k, params.slab.ptr, const_cast<float *>(host_slab), true);
    resetOptimizer();
}

void MultiFitter::downloadParams(int k, float *host_slab) {
    rowsCopy(k, params.slab.ptr, host_slab, false);
}

void MultiFitter::downloadGrads(int k, float *host_slab) {
    rowsCopy(k, grads.slab.ptr, host_slab, false);
}

void MultiFitter::downloadImage(int k, float *host_rgb) {
    if (k < 0 || k >= size()) throw std::runtime_error("[diffsplat] MultiFitter: bad instance");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const MultiFitInstance &q = inst[(size_t)k];
    hip_check(hipMemcpy(host_rgb, image.ptr + 3 * pix_base[(size_t)k],
                        sizeof(float) * 3 * (size_t)q.width * q.height, hipMemcpyDeviceToHost),
              "hipMemcpy");
}

void MultiFitter::resetOptimizer() {
    optimizer.reset(stream);
    iter_count = 0;
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void MultiFitter::forwardOnly() {
    ds_gauss3d pv = params.view();
    ds_splat2d sv = s2.view();
    check(ds_activ_project_fwd_multi(&pv, &sv, sh_degree, cams.ptr, splat_inst, stream),
          "MultiFitter::forward");
    check(ds_raster_fwd(ras, &sv, image.ptr, target.ptr, loss.ptr, grad_rgb.ptr, stream),
          "MultiFitter::forward");
}

void MultiFitter::forwardBackward() {
    ds_splat2d gv = gs2.view();
    check(ds_splat2d_zero(&gv, stream), "MultiFitter::zeroGrad");
    ds_gauss3d pv = params.view();
    ds_gauss3d gp = grads.view();
    ds_splat2d sv = s2.view();
    check(ds_activ_project_fwd_multi(&pv, &sv, sh_degree, cams.ptr, splat_inst, stream),
          "MultiFitter::forward");
    check(ds_raster_fwd_bwd(ras, &sv, image.ptr, target.ptr, loss.ptr, grad_rgb.ptr, &gv, stream),
          "MultiFitter::forward_backward");
    check(ds_activ_project_bwd_multi(&pv, &gv, &gp, sh_degree, cams.ptr, splat_inst, grad_mode,
                                     stream),
          "MultiFitter::backward");
}

// As ImageFitter's step: the rasterizer's first-kernel work (with the
// zeroing of the Splat2D gradients and the Adam step counter) in the
// activation/projection kernel and Adam in the projection backward, when the
// optimiser's groups are the slab's fields.
void MultiFitter::stepOnce() {
    std::vector<ds_adam_group> tab = optimizer.table();
    ds::AdamFields af;
    const ds_gauss3d pv = params.view();
    const ds_gauss3d gp = grads.view();
    if (pv.count <= 0 ||
        !adam_fields_of(tab, pv, gp, sh_degree, optimizer.getParams(), optimizer.stepDevice(),
                        af)) {
        forwardBackward();
        optimizer.step(stream);
        return;
    }
    ds_splat2d gv = gs2.view();
    ds_splat2d sv = s2.view();
    ds::ForwardStart fs;
    check(ds::raster_step_start(ras, pv.count, loss.ptr, gv, optimizer.stepDevice(), &fs),
          "MultiFitter::step");
    const float4 *ct = reinterpret_cast<const float4 *>(cams.ptr);
    check(ds::launch_activ_project_fwd_start_multi(pv, sv, sh_degree, ct, splat_inst, fs, stream),
          "MultiFitter::step");
    check(ds::raster_step_fwd_bwd(ras, sv, image.ptr, target.ptr, loss.ptr, grad_rgb.ptr, gv,
                                  optimizer.stepDevice(), stream, true),
          "MultiFitter::step");
    check(ds::launch_activ_project_bwd_adam_multi(pv, gv, gp, sh_degree, ct, splat_inst,
                                                  grad_mode, af, stream),
          "MultiFitter::step");
    optimizer.countStep();
}

void MultiFitter::step() {
    ++iter_count;
    if (!use_graph) {
        stepOnce();
        return;
    }
    if (!graph_exec) {
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        try {
            stepOnce();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0),
                  "hipGraphInstantiate");
    }
    hip_check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch");
}

void MultiFitter::losses(float *host) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    hip_check(hipMemcpy(host, loss.ptr, sizeof(float) * inst.size(), hipMemcpyDeviceToHost),
              "hipMemcpy");
}

void MultiFitter::autotune(int reps, int *best_fwd, int *best_bwd, float *ms_fwd,
                           float *ms_bwd) {
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    auto timed = [&](auto &&body) {
        body();  // warm-up
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        for (int k = 0; k < reps; ++k) body();
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / reps;
    };
    int fwd0 = 1, bwd0 = 1;
    check(ds_raster_get_tuning(ras, &fwd0, &bwd0), "autotune");
    const int opts[3] = {1, 2, 4};
    float tfw[3], tbw[3];
    for (int i = 0; i < 3; ++i) {
        check(ds_raster_set_tuning(ras, opts[i], bwd0), "autotune");
        tfw[i] = timed([&] { forwardOnly(); });
    }
    int bf = opts[pick_with_margin(tfw, fwd0)];
    check(ds_raster_set_tuning(ras, bf, bwd0), "autotune");
    forwardOnly();
    ds_splat2d sv = s2.view(), gv = gs2.view();
    for (int i = 0; i < 3; ++i) {
        check(ds_raster_set_tuning(ras, bf, opts[i]), "autotune");
        tbw[i] = timed([&] { check(ds_raster_bwd(ras, &sv, grad_rgb.ptr, &gv, stream), "bwd"); });
    }
    const int ib = pick_with_margin(tbw, bwd0);
    int bb = opts[ib];
    const float tf = tfw[bf == 1 ? 0 : bf == 2 ? 1 : 2], tb = tbw[ib];
    check(ds_raster_set_tuning(ras, bf, bb), "autotune");
    // the step's compositing: one fused kernel when the shapes are equal
    ds_gauss3d pv = params.view();
    auto fwd_bwd = [&] {
        return timed([&] {
            check(ds_activ_project_fwd_multi(&pv, &sv, sh_degree, cams.ptr, splat_inst, stream),
                  "forward");
            check(ds_raster_fwd_bwd(ras, &sv, image.ptr, target.ptr, loss.ptr, grad_rgb.ptr, &gv,
                                    stream),
                  "forward_backward");
        });
    };
    tune_fused_shapes(ras, bf, bb, fwd_bwd);
    tune_tile_order(ras, fwd_bwd);
    tune_depth_sort(ras, sv, stream, reps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dropGraph();
    if (best_fwd) *best_fwd = bf;
    if (best_bwd) *best_bwd = bb;
    if (ms_fwd) *ms_fwd = tf;
    if (ms_bwd) *ms_bwd = tb;
}

// ============================================================ SplatRenderer

SplatRenderer::SplatRenderer() = default;

SplatRenderer::~SplatRenderer() {
    if (stream) (void)hipStreamDestroy(stream);
}

void SplatRenderer::init(int width, int height, int count, int deg, int64_t isect_cap) {
    if (width <= 0 || height <= 0 || count < 0 || deg < 0 || deg > 3)
        throw std::runtime_error("[diffsplat] SplatRenderer::init: bad configuration");
    w = width;
    h = height;
    sh_degree = deg;
    if (!stream) hip_check(hipStreamCreate(&stream), "hipStreamCreate");
    gaussian_params.allocate(count, deg);
    gp_layer.setStream(stream);
    ras_layer.setStream(stream);
    gp_layer.setSHDegree(deg);
    gp_layer.allocate(count);
    ras_layer.allocate(width, height, count, isect_cap);
    ras_layer.releaseBackward();  // frames only
    gp_layer.setInput(&gaussian_params);
    ras_layer.setInput(&gp_layer.getOutput());
    pipeline.clear();
    pipeline.add(&gp_layer);
    pipeline.add(&ras_layer);
}

void SplatRenderer::uploadParams(const float *host_slab) {
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    gaussian_params.slab.upload(host_slab, gaussian_params.slab.size);
}

void SplatRenderer::render(const float view[16], const float proj[16], const float cam_pos[3]) {
    gp_layer.setCameraPosition(cam_pos[0], cam_pos[1], cam_pos[2]);
    gp_layer.setCamera(view, proj);
    pipeline.forward();
}

int SplatRenderer::autotune(int reps, float *ms_best) {
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    ds_raster *r = ras_layer.raster();
    int fwd0 = 1, bwd0 = 1;
    check(ds_raster_get_tuning(r, &fwd0, &bwd0), "autotune");
    const int opts[3] = {1, 2, 4};
    float t[3];
    for (int i = 0; i < 3; ++i) {
        check(ds_raster_set_tuning(r, opts[i], bwd0), "autotune");
        pipeline.forward();  // warm-up
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        for (int k = 0; k < reps; ++k) pipeline.forward();
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        t[i] = ms / reps;
    }
    const int ib = pick_with_margin(t, fwd0);
    check(ds_raster_set_tuning(r, opts[ib], bwd0), "autotune");
    tune_tile_order(r, [&] {
        pipeline.forward();  // warm-up
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        for (int k = 0; k < reps; ++k) pipeline.forward();
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / reps;
    });
    pipeline.forward();
    tune_depth_sort(r, gp_layer.getOutput().view(), stream, reps);
    pipeline.forward();  // the frame again, with the kept sort
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_best) *ms_best = t[ib];
    return opts[ib];
}

}  // namespace dsh
