// C-ABI entry points: runtime helpers, layer kernels, MSE, Adam, and the
// fitter / renderer objects (wrapping the host C++ classes).
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "ds_internal.hpp"
#include "host/pipelines.hpp"

#ifndef DS_VERSION
#define DS_VERSION "0.1.0"
#endif

namespace ds {

static thread_local std::string g_last_error;

ds_status set_error(ds_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}
void clear_error() { g_last_error.clear(); }

}  // namespace ds

using namespace ds;

#define DS_GUARD_BEGIN try {
#define DS_GUARD_END                                                       \
    }                                                                      \
    catch (const std::exception &e) {                                      \
        return set_error(DS_ERR_HIP, e.what());                            \
    }                                                                      \
    catch (...) {                                                          \
        return set_error(DS_ERR_HIP, "unknown C++ exception");             \
    }

extern "C" {

const char *ds_status_string(ds_status s) {
    switch (s) {
        case DS_OK: return "DS_OK";
        case DS_ERR_INVALID_ARG: return "DS_ERR_INVALID_ARG";
        case DS_ERR_HIP: return "DS_ERR_HIP";
        case DS_ERR_OUT_OF_MEMORY: return "DS_ERR_OUT_OF_MEMORY";
        case DS_ERR_CAPACITY: return "DS_ERR_CAPACITY";
        case DS_ERR_STATE: return "DS_ERR_STATE";
    }
    return "DS_ERR_UNKNOWN";
}

const char *ds_last_error(void) { return g_last_error.c_str(); }

const char *ds_version(void) { return "diffsplat_amd " DS_VERSION " (gfx950)"; }

int ds_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

ds_status ds_device_name(char *buf, size_t len) {
    clear_error();
    DS_REQUIRE(buf && len > 0, "NULL buffer");
    buf[0] = 0;
    if (ds_device_count() == 0) return DS_OK;
    hipDeviceProp_t p;
    DS_TRY_HIP(hipGetDeviceProperties(&p, 0));
    std::strncpy(buf, p.gcnArchName[0] ? p.gcnArchName : p.name, len - 1);
    buf[len - 1] = 0;
    return DS_OK;
}

ds_status ds_malloc(void **dptr, size_t bytes) {
    clear_error();
    DS_REQUIRE(dptr, "NULL dptr");
    *dptr = nullptr;
    if (bytes == 0) return DS_OK;
    DS_TRY_HIP(hipMalloc(dptr, bytes));
    return DS_OK;
}

ds_status ds_free(void *dptr) {
    clear_error();
    if (dptr) DS_TRY_HIP(hipFree(dptr));
    return DS_OK;
}

ds_status ds_memcpy_htod(void *dst, const void *src, size_t bytes) {
    clear_error();
    if (bytes == 0) return DS_OK;
    DS_REQUIRE(dst && src, "NULL pointer");
    DS_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DS_OK;
}

ds_status ds_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    clear_error();
    if (bytes == 0) return DS_OK;
    DS_REQUIRE(dst && src, "NULL pointer");
    DS_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DS_OK;
}

ds_status ds_memcpy_dtod(void *dst, const void *src, size_t bytes, ds_stream_t s) {
    clear_error();
    if (bytes == 0) return DS_OK;
    DS_REQUIRE(dst && src, "NULL pointer");
    DS_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
    return DS_OK;
}

ds_status ds_fill_u32(void *dptr, uint32_t value, size_t count, ds_stream_t s) {
    clear_error();
    if (count == 0) return DS_OK;
    DS_REQUIRE(dptr, "NULL pointer");
    return fill_u32(dptr, value, count, as_stream(s));
}

ds_status ds_synchronize(void) {
    clear_error();
    DS_TRY_HIP(hipDeviceSynchronize());
    return DS_OK;
}

ds_status ds_stream_create(ds_stream_t *s) {
    clear_error();
    DS_REQUIRE(s, "NULL stream pointer");
    hipStream_t st = nullptr;
    DS_TRY_HIP(hipStreamCreate(&st));
    *s = st;
    return DS_OK;
}

ds_status ds_stream_destroy(ds_stream_t s) {
    clear_error();
    if (s) DS_TRY_HIP(hipStreamDestroy(as_stream(s)));
    return DS_OK;
}

ds_status ds_stream_synchronize(ds_stream_t s) {
    clear_error();
    DS_TRY_HIP(hipStreamSynchronize(as_stream(s)));
    return DS_OK;
}

ds_status ds_event_create(ds_event_t *e) {
    clear_error();
    DS_REQUIRE(e, "NULL event pointer");
    hipEvent_t ev = nullptr;
    DS_TRY_HIP(hipEventCreate(&ev));
    *e = ev;
    return DS_OK;
}

ds_status ds_event_destroy(ds_event_t e) {
    clear_error();
    if (e) DS_TRY_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)));
    return DS_OK;
}

ds_status ds_event_record(ds_event_t e, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(e, "NULL event");
    DS_TRY_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(e), as_stream(s)));
    return DS_OK;
}

ds_status ds_event_synchronize(ds_event_t e) {
    clear_error();
    DS_REQUIRE(e, "NULL event");
    DS_TRY_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)));
    return DS_OK;
}

ds_status ds_event_elapsed_ms(ds_event_t a, ds_event_t b, float *ms) {
    clear_error();
    DS_REQUIRE(a && b && ms, "NULL argument");
    DS_TRY_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(a),
                                   reinterpret_cast<hipEvent_t>(b)));
    return DS_OK;
}

// ------------------------------------------------------------------ layers

static ds_status check_gauss(const ds_gauss3d *g, int deg, const char *what) {
    DS_REQUIRE(g, std::string(what) + ": NULL view");
    DS_REQUIRE(deg >= 0 && deg <= 3, std::string(what) + ": sh_degree must be in [0,3]");
    DS_REQUIRE(g->count >= 0, std::string(what) + ": negative count");
    if (g->count == 0) return DS_OK;
    DS_REQUIRE(g->pos_x && g->pos_y && g->pos_z && g->scale_x && g->scale_y && g->scale_z &&
                   g->rot_w && g->rot_x && g->rot_y && g->rot_z && g->sh_dc_r && g->sh_dc_g &&
                   g->sh_dc_b && g->logit_opacity,
               std::string(what) + ": NULL field");
    if (deg > 0) {
        DS_REQUIRE(g->sh_num_bands >= ds_sh_bands(deg),
                   std::string(what) + ": sh_num_bands smaller than the SH degree needs");
        DS_REQUIRE(g->sh_rest_r && g->sh_rest_g && g->sh_rest_b,
                   std::string(what) + ": NULL sh_rest field");
    }
    return DS_OK;
}

static ds_status check_mode(int m) {
    DS_REQUIRE(m == DS_GRAD_EXACT || m == DS_GRAD_REFERENCE,
               "grad_mode must be DS_GRAD_EXACT or DS_GRAD_REFERENCE");
    return DS_OK;
}

static Cam3 cam_of(const float *c) { return c ? Cam3{c[0], c[1], c[2]} : Cam3{0.f, 0.f, 0.f}; }

static PV pv_of(const float *m) {
    PV p;
    std::memcpy(p.m, m, sizeof(p.m));
    return p;
}

ds_status ds_activ_fwd(const ds_gauss3d *in, const ds_splat3d *out, int deg,
                       const float cam_pos[3], ds_stream_t s) {
    clear_error();
    DS_TRY(check_gauss(in, deg, "ds_activ_fwd"));
    DS_REQUIRE(out && out->count >= in->count, "ds_activ_fwd: output view too small");
    return launch_activ_fwd(*in, *out, deg, cam_of(cam_pos), as_stream(s));
}

ds_status ds_activ_bwd(const ds_gauss3d *in, const ds_splat3d *out, const ds_splat3d *gout,
                       const ds_gauss3d *gin, int deg, const float cam_pos[3], int grad_mode,
                       ds_stream_t s) {
    clear_error();
    DS_TRY(check_gauss(in, deg, "ds_activ_bwd"));
    DS_TRY(check_gauss(gin, deg, "ds_activ_bwd grad_in"));
    DS_REQUIRE(out && gout, "ds_activ_bwd: NULL view");
    DS_TRY(check_mode(grad_mode));
    return launch_activ_bwd(*in, *out, *gout, *gin, deg, cam_of(cam_pos), grad_mode,
                            as_stream(s));
}

ds_status ds_project_fwd(const ds_splat3d *in, const ds_splat2d *out, const float pv[16],
                         ds_stream_t s) {
    clear_error();
    DS_REQUIRE(in && out && pv, "ds_project_fwd: NULL argument");
    return launch_project_fwd(*in, *out, pv_of(pv), as_stream(s));
}

ds_status ds_project_bwd(const ds_splat3d *in, const ds_splat2d *gout, const ds_splat3d *gin,
                         const float pv[16], int grad_mode, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(in && gout && gin && pv, "ds_project_bwd: NULL argument");
    DS_TRY(check_mode(grad_mode));
    return launch_project_bwd(*in, *gout, *gin, pv_of(pv), grad_mode, as_stream(s));
}

ds_status ds_activ_project_fwd(const ds_gauss3d *in, const ds_splat2d *out, int deg,
                               const float cam_pos[3], const float pv[16], ds_stream_t s) {
    clear_error();
    DS_TRY(check_gauss(in, deg, "ds_activ_project_fwd"));
    DS_REQUIRE(out && pv, "ds_activ_project_fwd: NULL argument");
    return launch_activ_project_fwd(*in, *out, deg, cam_of(cam_pos), pv_of(pv), as_stream(s));
}

ds_status ds_activ_project_bwd(const ds_gauss3d *in, const ds_splat2d *gout,
                               const ds_gauss3d *gin, int deg, const float cam_pos[3],
                               const float pv[16], int grad_mode, ds_stream_t s) {
    clear_error();
    DS_TRY(check_mode(grad_mode));
    DS_TRY(check_gauss(in, deg, "ds_activ_project_bwd"));
    DS_TRY(check_gauss(gin, deg, "ds_activ_project_bwd grad_in"));
    DS_REQUIRE(gout && pv, "ds_activ_project_bwd: NULL argument");
    return launch_activ_project_bwd(*in, *gout, *gin, deg, cam_of(cam_pos), pv_of(pv),
                                    grad_mode, as_stream(s));
}

ds_status ds_activ_project_fwd_multi(const ds_gauss3d *in, const ds_splat2d *out, int deg,
                                     const float *cam_tab, const uint16_t *splat_inst,
                                     ds_stream_t s) {
    clear_error();
    DS_TRY(check_gauss(in, deg, "ds_activ_project_fwd_multi"));
    DS_REQUIRE(out && cam_tab && (splat_inst || in->count == 0),
               "ds_activ_project_fwd_multi: NULL argument");
    return launch_activ_project_fwd_multi(*in, *out, deg,
                                          reinterpret_cast<const float4 *>(cam_tab), splat_inst,
                                          as_stream(s));
}

ds_status ds_activ_project_bwd_multi(const ds_gauss3d *in, const ds_splat2d *gout,
                                     const ds_gauss3d *gin, int deg, const float *cam_tab,
                                     const uint16_t *splat_inst, int grad_mode, ds_stream_t s) {
    clear_error();
    DS_TRY(check_mode(grad_mode));
    DS_TRY(check_gauss(in, deg, "ds_activ_project_bwd_multi"));
    DS_TRY(check_gauss(gin, deg, "ds_activ_project_bwd_multi grad_in"));
    DS_REQUIRE(gout && cam_tab && (splat_inst || in->count == 0),
               "ds_activ_project_bwd_multi: NULL argument");
    return launch_activ_project_bwd_multi(*in, *gout, *gin, deg,
                                          reinterpret_cast<const float4 *>(cam_tab), splat_inst,
                                          grad_mode, as_stream(s));
}

ds_status ds_mse_fwd(const float *pixels, const float *target, int64_t num_pixels,
                     float *loss_dev, ds_stream_t s) {
    clear_error();
    DS_REQUIRE(loss_dev && num_pixels >= 0, "ds_mse_fwd: bad argument");
    DS_REQUIRE(num_pixels == 0 || (pixels && target), "ds_mse_fwd: NULL image");
    return launch_mse_fwd(pixels, target, num_pixels, loss_dev, as_stream(s));
}

ds_status ds_mse_bwd(const float *pixels, const float *target, int64_t num_pixels, float *grad,
                     ds_stream_t s) {
    clear_error();
    DS_REQUIRE(num_pixels >= 0, "ds_mse_bwd: bad argument");
    DS_REQUIRE(num_pixels == 0 || (pixels && target && grad), "ds_mse_bwd: NULL pointer");
    return launch_mse_bwd(pixels, target, num_pixels, grad, as_stream(s));
}

static ds_status check_adam(const ds_adam_group *g, int n, const ds_adam_hparams *hp) {
    DS_REQUIRE(hp, "ds_adam_step: NULL hparams");
    DS_REQUIRE(n >= 0 && n <= DS_ADAM_MAX_GROUPS, "ds_adam_step: n_groups out of range");
    DS_REQUIRE(n == 0 || g, "ds_adam_step: NULL groups");
    for (int k = 0; k < n; ++k) {
        DS_REQUIRE(g[k].n >= 0, "ds_adam_step: negative group size");
        DS_REQUIRE(g[k].n == 0 || (g[k].param && g[k].grad && g[k].m && g[k].v),
                   "ds_adam_step: NULL group pointer");
    }
    return DS_OK;
}

ds_status ds_adam_step(const ds_adam_group *groups, int n_groups, const ds_adam_hparams *hp,
                       int32_t step, ds_stream_t s) {
    clear_error();
    DS_TRY(check_adam(groups, n_groups, hp));
    DS_REQUIRE(step >= 1, "ds_adam_step: step must be >= 1");
    return launch_adam(groups, n_groups, *hp, step, nullptr, as_stream(s));
}

ds_status ds_adam_step_dev(const ds_adam_group *groups, int n_groups, const ds_adam_hparams *hp,
                           int32_t *step_dev, ds_stream_t s) {
    clear_error();
    DS_TRY(check_adam(groups, n_groups, hp));
    DS_REQUIRE(step_dev, "ds_adam_step_dev: NULL step counter");
    return launch_adam(groups, n_groups, *hp, 0, step_dev, as_stream(s));
}

// ------------------------------------------------------------------ fitter

struct ds_fitter {
    dsh::ImageFitter impl;
};

void ds_fitter_default_config(int width, int height, int count, ds_fitter_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->width = width;
    cfg->height = height;
    cfg->count = count;
    cfg->sh_degree = 0;
    dsh::ImageFitter::referenceCamera(width, height, cfg->view, cfg->proj);
    cfg->cam_pos[0] = 0.f;
    cfg->cam_pos[1] = 0.f;
    cfg->cam_pos[2] = 5.f;
    dsh::Gaussian3DAdamConfig d;
    cfg->lr_pos = d.lr_pos;
    cfg->lr_scale = d.lr_scale;
    cfg->lr_rot = d.lr_rot;
    cfg->lr_sh_dc = d.lr_sh_dc;
    cfg->lr_sh_rest = d.lr_sh_rest;
    cfg->lr_opacity = d.lr_opacity;
    cfg->beta1 = d.adam.beta1;
    cfg->beta2 = d.adam.beta2;
    cfg->epsilon = d.adam.epsilon;
    cfg->isect_cap = 0;
    cfg->fused = 1;
    cfg->grad_mode = DS_GRAD_EXACT;
    cfg->lr_pos_decay = 1.f;
}

ds_status ds_fitter_create(const ds_fitter_config *cfg, ds_fitter **out) {
    clear_error();
    DS_REQUIRE(cfg && out, "NULL argument");
    DS_REQUIRE(cfg->width > 0 && cfg->height > 0 && cfg->count >= 0, "bad fitter size");
    DS_REQUIRE(cfg->sh_degree >= 0 && cfg->sh_degree <= 3, "sh_degree must be in [0,3]");
    DS_GUARD_BEGIN
    auto f = std::make_unique<ds_fitter>();
    dsh::FitterOptions o;
    o.width = cfg->width;
    o.height = cfg->height;
    o.count = cfg->count;
    o.sh_degree = cfg->sh_degree;
    std::memcpy(o.view, cfg->view, sizeof(o.view));
    std::memcpy(o.proj, cfg->proj, sizeof(o.proj));
    std::memcpy(o.cam_pos, cfg->cam_pos, sizeof(o.cam_pos));
    o.lr.lr_pos = cfg->lr_pos;
    o.lr.lr_scale = cfg->lr_scale;
    o.lr.lr_rot = cfg->lr_rot;
    o.lr.lr_sh_dc = cfg->lr_sh_dc;
    o.lr.lr_sh_rest = cfg->lr_sh_rest;
    o.lr.lr_opacity = cfg->lr_opacity;
    o.lr.lr_pos_decay = cfg->lr_pos_decay > 0.f ? cfg->lr_pos_decay : 1.f;
    o.lr.adam.beta1 = cfg->beta1;
    o.lr.adam.beta2 = cfg->beta2;
    o.lr.adam.epsilon = cfg->epsilon;
    o.isect_cap = cfg->isect_cap;
    o.fused = cfg->fused != 0;
    DS_REQUIRE(cfg->grad_mode == DS_GRAD_EXACT || cfg->grad_mode == DS_GRAD_REFERENCE,
               "grad_mode must be DS_GRAD_EXACT or DS_GRAD_REFERENCE");
    o.grad_mode = cfg->grad_mode;
    f->impl.init(o);
    f->impl.initLayers();
    *out = f.release();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_destroy(ds_fitter *f) {
    clear_error();
    DS_GUARD_BEGIN
    delete f;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_set_target(ds_fitter *f, const float *host_rgb) {
    clear_error();
    DS_REQUIRE(f && host_rgb, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.setTarget(host_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_upload_params(ds_fitter *f, const float *host_slab) {
    clear_error();
    DS_REQUIRE(f && host_slab, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.uploadParams(host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_download_params(ds_fitter *f, float *host_slab) {
    clear_error();
    DS_REQUIRE(f && host_slab, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.downloadParams(host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_download_grads(ds_fitter *f, float *host_slab) {
    clear_error();
    DS_REQUIRE(f && host_slab, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.downloadGrads(host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_download_moments(ds_fitter *f, float *host_m, float *host_v) {
    clear_error();
    DS_REQUIRE(f && (host_m || host_v), "NULL argument");
    DS_GUARD_BEGIN
    f->impl.downloadMoments(host_m, host_v);
    return DS_OK;
    DS_GUARD_END
}

void ds_densify_default_config(ds_densify_config *cfg) {
    if (!cfg) return;
    cfg->grad_threshold = 2e-4f;
    cfg->percent_dense = 0.01f;
    cfg->scene_extent = 0.f;
    cfg->min_opacity = 0.005f;
    cfg->max_scale_frac = 0.f;
    cfg->split_scale_div = 1.6f;
    cfg->seed = 0;
}

ds_status ds_fitter_set_densify_stats(ds_fitter *f, int enable) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.setDensifyStats(enable != 0);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_densify_stats(ds_fitter *f, float *host_grad_sum, uint32_t *host_visits) {
    clear_error();
    DS_REQUIRE(f && (host_grad_sum || host_visits), "NULL argument");
    DS_GUARD_BEGIN
    f->impl.downloadDensifyStats(host_grad_sum, host_visits);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_densify(ds_fitter *f, const ds_densify_config *cfg,
                            ds_densify_result *result, uint8_t *host_actions) {
    clear_error();
    DS_REQUIRE(f && cfg, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.densify(*cfg, result, host_actions);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_count(ds_fitter *f, int32_t *count) {
    clear_error();
    DS_REQUIRE(f && count, "NULL argument");
    *count = f->impl.count();
    return DS_OK;
}

ds_status ds_fitter_step(ds_fitter *f, int n_steps) {
    clear_error();
    DS_REQUIRE(f && n_steps >= 0, "bad argument");
    DS_GUARD_BEGIN
    f->impl.steps(n_steps);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_prepare(ds_fitter *f, int n_steps) {
    clear_error();
    DS_REQUIRE(f && n_steps >= 0, "bad argument");
    DS_GUARD_BEGIN
    f->impl.prepare(n_steps);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_forward(ds_fitter *f) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.forwardOnly();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_loss(ds_fitter *f, float *loss) {
    clear_error();
    DS_REQUIRE(f && loss, "NULL argument");
    DS_GUARD_BEGIN
    *loss = f->impl.getLoss();
    return ds::raster_take_error(f->impl.raster(), f->impl.getStream());
    DS_GUARD_END
}

ds_status ds_fitter_synchronize(ds_fitter *f) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_TRY_HIP(hipStreamSynchronize(f->impl.getStream()));
    return ds::raster_take_error(f->impl.raster(), f->impl.getStream());
}

ds_status ds_fitter_download_image(ds_fitter *f, float *host_rgb) {
    clear_error();
    DS_REQUIRE(f && host_rgb, "NULL argument");
    DS_TRY_HIP(hipStreamSynchronize(f->impl.getStream()));
    const size_t bytes = (size_t)f->impl.width() * f->impl.height() * 3 * sizeof(float);
    DS_TRY_HIP(hipMemcpy(host_rgb, f->impl.getOutput(), bytes, hipMemcpyDeviceToHost));
    return ds::raster_take_error(f->impl.raster(), f->impl.getStream());
}

ds_status ds_fitter_get_raster(ds_fitter *f, ds_raster **r) {
    clear_error();
    DS_REQUIRE(f && r, "NULL argument");
    *r = f->impl.raster();
    return DS_OK;
}

ds_status ds_fitter_get_stream(ds_fitter *f, ds_stream_t *s) {
    clear_error();
    DS_REQUIRE(f && s, "NULL argument");
    *s = f->impl.getStream();
    return DS_OK;
}

ds_status ds_fitter_set_graph(ds_fitter *f, int enable) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    f->impl.setUseGraph(enable != 0);
    return DS_OK;
}

ds_status ds_fitter_autotune(ds_fitter *f, int reps, int *fwd_ppt, int *bwd_ppt, float *fwd_ms,
                             float *bwd_ms) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.autotune(reps, fwd_ppt, bwd_ppt, fwd_ms, bwd_ms);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_step_count(ds_fitter *f, int32_t *step) {
    clear_error();
    DS_REQUIRE(f && step, "NULL argument");
    *step = f->impl.getIterCount();
    return DS_OK;
}

ds_status ds_fitter_views(ds_fitter *f, ds_gauss3d *params, ds_splat2d *splat2d,
                          ds_splat2d *grad_splat2d, float **grad_rgb) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.views(params, splat2d, grad_splat2d, grad_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_save_ply(ds_fitter *f, const char *path) {
    clear_error();
    DS_REQUIRE(f && path, "NULL argument");
    DS_GUARD_BEGIN
    const int n = f->impl.count(), deg = f->impl.shDegree();
    DS_REQUIRE(n > 0, "the fitter holds no splats");
    std::vector<float> slab((size_t)ds_slab_floats(n, deg));
    f->impl.downloadParams(slab.data());
    return ds_ply_save(path, slab.data(), n, deg);
    DS_GUARD_END
}

ds_status ds_fitter_forward_backward(ds_fitter *f) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.forwardBackward();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_param_slabs(ds_fitter *f, float **params_dev, const float **grads_dev) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    if (params_dev) *params_dev = f->impl.paramSlab();
    if (grads_dev) *grads_dev = f->impl.gradSlab();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_set_target_device(ds_fitter *f, const float *dev_rgb) {
    clear_error();
    DS_REQUIRE(f && dev_rgb, "NULL argument");
    DS_GUARD_BEGIN
    f->impl.setTargetDevice(dev_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_fitter_reset_optimizer(ds_fitter *f) {
    clear_error();
    DS_REQUIRE(f, "NULL fitter");
    DS_GUARD_BEGIN
    f->impl.resetOptimizer();
    return DS_OK;
    DS_GUARD_END
}

// ------------------------------------------------------------------ batch

// ------------------------------------------------------------------ multifit

struct ds_multifit {
    dsh::MultiFitter impl;
};

ds_status ds_multifit_create(const ds_multifit_instance *inst, int n_inst,
                             const ds_fitter_config *cfg, int64_t isect_cap, ds_multifit **out) {
    clear_error();
    DS_REQUIRE(inst && cfg && out, "NULL argument");
    DS_REQUIRE(n_inst >= 1 && n_inst <= 65535, "n_inst must be in [1, 65535]");
    DS_REQUIRE(cfg->sh_degree >= 0 && cfg->sh_degree <= 3, "sh_degree must be in [0,3]");
    DS_REQUIRE(cfg->grad_mode == DS_GRAD_EXACT || cfg->grad_mode == DS_GRAD_REFERENCE,
               "grad_mode must be DS_GRAD_EXACT or DS_GRAD_REFERENCE");
    for (int k = 0; k < n_inst; ++k)
        DS_REQUIRE(inst[k].width > 0 && inst[k].height > 0 && inst[k].count >= 0,
                   "bad instance size");
    DS_GUARD_BEGIN
    auto m = std::make_unique<ds_multifit>();
    std::vector<dsh::MultiFitInstance> v((size_t)n_inst);
    for (int k = 0; k < n_inst; ++k) {
        dsh::MultiFitInstance &q = v[(size_t)k];
        q.width = inst[k].width;
        q.height = inst[k].height;
        q.count = inst[k].count;
        std::memcpy(q.view, inst[k].view, sizeof(q.view));
        std::memcpy(q.proj, inst[k].proj, sizeof(q.proj));
        std::memcpy(q.cam_pos, inst[k].cam_pos, sizeof(q.cam_pos));
    }
    dsh::Gaussian3DAdamConfig lr;
    lr.lr_pos = cfg->lr_pos;
    lr.lr_scale = cfg->lr_scale;
    lr.lr_rot = cfg->lr_rot;
    lr.lr_sh_dc = cfg->lr_sh_dc;
    lr.lr_sh_rest = cfg->lr_sh_rest;
    lr.lr_opacity = cfg->lr_opacity;
    lr.lr_pos_decay = cfg->lr_pos_decay > 0.f ? cfg->lr_pos_decay : 1.f;
    lr.adam.beta1 = cfg->beta1;
    lr.adam.beta2 = cfg->beta2;
    lr.adam.epsilon = cfg->epsilon;
    m->impl.init(v, cfg->sh_degree, lr, isect_cap, cfg->grad_mode);
    *out = m.release();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_destroy(ds_multifit *m) {
    clear_error();
    DS_GUARD_BEGIN
    delete m;
    return DS_OK;
    DS_GUARD_END
}

#define DS_MF_INST(m, k, p)                                                   \
    DS_REQUIRE((m) && (p), "NULL argument");                                  \
    DS_REQUIRE((k) >= 0 && (k) < (m)->impl.size(), "instance index out of range")

ds_status ds_multifit_set_target(ds_multifit *m, int k, const float *host_rgb) {
    clear_error();
    DS_MF_INST(m, k, host_rgb);
    DS_GUARD_BEGIN
    m->impl.setTarget(k, host_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_upload_params(ds_multifit *m, int k, const float *host_slab) {
    clear_error();
    DS_MF_INST(m, k, host_slab);
    DS_GUARD_BEGIN
    m->impl.uploadParams(k, host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_download_params(ds_multifit *m, int k, float *host_slab) {
    clear_error();
    DS_MF_INST(m, k, host_slab);
    DS_GUARD_BEGIN
    m->impl.downloadParams(k, host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_download_grads(ds_multifit *m, int k, float *host_slab) {
    clear_error();
    DS_MF_INST(m, k, host_slab);
    DS_GUARD_BEGIN
    m->impl.downloadGrads(k, host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_download_image(ds_multifit *m, int k, float *host_rgb) {
    clear_error();
    DS_MF_INST(m, k, host_rgb);
    DS_GUARD_BEGIN
    m->impl.downloadImage(k, host_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_step(ds_multifit *m, int n_steps) {
    clear_error();
    DS_REQUIRE(m && n_steps >= 0, "bad argument");
    DS_GUARD_BEGIN
    for (int k = 0; k < n_steps; ++k) m->impl.step();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_forward(ds_multifit *m) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    DS_GUARD_BEGIN
    m->impl.forwardOnly();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_forward_backward(ds_multifit *m) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    DS_GUARD_BEGIN
    m->impl.forwardBackward();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_losses(ds_multifit *m, float *losses) {
    clear_error();
    DS_REQUIRE(m && losses, "NULL argument");
    DS_GUARD_BEGIN
    m->impl.losses(losses);
    return ds::raster_take_error(m->impl.raster(), m->impl.getStream());
    DS_GUARD_END
}

ds_status ds_multifit_synchronize(ds_multifit *m) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    DS_TRY_HIP(hipStreamSynchronize(m->impl.getStream()));
    return ds::raster_take_error(m->impl.raster(), m->impl.getStream());
}

ds_status ds_multifit_set_graph(ds_multifit *m, int enable) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    m->impl.setUseGraph(enable != 0);
    return DS_OK;
}

ds_status ds_multifit_autotune(ds_multifit *m, int reps, int *fwd_ppt, int *bwd_ppt,
                               float *fwd_ms, float *bwd_ms) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    DS_GUARD_BEGIN
    m->impl.autotune(reps, fwd_ppt, bwd_ppt, fwd_ms, bwd_ms);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_step_count(ds_multifit *m, int32_t *step) {
    clear_error();
    DS_REQUIRE(m && step, "NULL argument");
    *step = m->impl.getIterCount();
    return DS_OK;
}

ds_status ds_multifit_get_raster(ds_multifit *m, ds_raster **r) {
    clear_error();
    DS_REQUIRE(m && r, "NULL argument");
    *r = m->impl.raster();
    return DS_OK;
}

ds_status ds_multifit_views(ds_multifit *m, ds_gauss3d *params, ds_splat2d *splat2d,
                            ds_splat2d *grad_splat2d, float **grad_rgb) {
    clear_error();
    DS_REQUIRE(m, "NULL multifit");
    DS_GUARD_BEGIN
    m->impl.views(params, splat2d, grad_splat2d, grad_rgb);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_multifit_get_stream(ds_multifit *m, ds_stream_t *s) {
    clear_error();
    DS_REQUIRE(m && s, "NULL argument");
    *s = m->impl.getStream();
    return DS_OK;
}

struct ds_batch {
    dsh::BatchFitter impl;
};

ds_status ds_batch_create(ds_fitter *const *fitters, int n, ds_batch **out) {
    clear_error();
    DS_REQUIRE(fitters && out, "NULL argument");
    DS_REQUIRE(n >= 1, "a batch needs at least one fitter");
    for (int i = 0; i < n; ++i) DS_REQUIRE(fitters[i], "NULL fitter in the batch");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) DS_REQUIRE(fitters[i] != fitters[j], "fitter listed twice");
    DS_GUARD_BEGIN
    auto b = std::make_unique<ds_batch>();
    std::vector<dsh::ImageFitter *> v((size_t)n);
    for (int i = 0; i < n; ++i) v[(size_t)i] = &fitters[i]->impl;
    b->impl.init(std::move(v));
    *out = b.release();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_batch_destroy(ds_batch *b) {
    clear_error();
    DS_GUARD_BEGIN
    delete b;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_batch_step(ds_batch *b, int n_steps) {
    clear_error();
    DS_REQUIRE(b, "NULL batch");
    DS_REQUIRE(n_steps >= 0, "n_steps < 0");
    DS_GUARD_BEGIN
    b->impl.step(n_steps);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_batch_synchronize(ds_batch *b) {
    clear_error();
    DS_REQUIRE(b, "NULL batch");
    DS_GUARD_BEGIN
    b->impl.synchronize();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_batch_set_streams(ds_batch *b, int n_streams) {
    clear_error();
    DS_REQUIRE(b, "NULL batch");
    DS_REQUIRE(n_streams >= 1 && n_streams <= 64, "n_streams must be in [1, 64]");
    DS_GUARD_BEGIN
    b->impl.setStreams(n_streams);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_batch_autotune(ds_batch *b, int reps, int *n_streams, float *ms_per_batch_step) {
    clear_error();
    DS_REQUIRE(b, "NULL batch");
    DS_GUARD_BEGIN
    const int ns = b->impl.autotune(reps, ms_per_batch_step);
    if (n_streams) *n_streams = ns;
    return DS_OK;
    DS_GUARD_END
}

// ------------------------------------------------------------------ renderer

struct ds_renderer {
    dsh::SplatRenderer impl;
    bool rendered = false;  // a camera has been set
};

ds_status ds_renderer_create(int width, int height, int count, int sh_degree, int64_t isect_cap,
                             ds_renderer **out) {
    clear_error();
    DS_REQUIRE(out, "NULL argument");
    DS_REQUIRE(width > 0 && height > 0 && count >= 0, "bad renderer size");
    DS_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "sh_degree must be in [0,3]");
    DS_GUARD_BEGIN
    auto r = std::make_unique<ds_renderer>();
    r->impl.init(width, height, count, sh_degree, isect_cap);
    *out = r.release();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_load_ply(const char *path, int width, int height, float scene_scale,
                               int64_t isect_cap, ds_renderer **out, int64_t *count,
                               int32_t *sh_degree) {
    clear_error();
    DS_REQUIRE(path && out, "NULL argument");
    int64_t n = 0;
    int32_t deg = 0;
    DS_TRY(ds_ply_info(path, &n, &deg));
    DS_REQUIRE(n <= (int64_t)INT32_MAX, "PLY has more splats than a renderer holds");
    DS_GUARD_BEGIN
    std::vector<float> slab((size_t)ds_slab_floats((int)n, deg));
    DS_TRY(ds_ply_load(path, slab.data(), n, deg));
    DS_TRY(ds_normalize_scene(slab.data(), n, scene_scale));
    auto r = std::make_unique<ds_renderer>();
    r->impl.init(width, height, (int)n, deg, isect_cap);
    r->impl.uploadParams(slab.data());
    *out = r.release();
    if (count) *count = n;
    if (sh_degree) *sh_degree = deg;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_destroy(ds_renderer *r) {
    clear_error();
    DS_GUARD_BEGIN
    delete r;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_upload_params(ds_renderer *r, const float *host_slab) {
    clear_error();
    DS_REQUIRE(r && host_slab, "NULL argument");
    DS_GUARD_BEGIN
    r->impl.uploadParams(host_slab);
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_render(ds_renderer *r, const float view[16], const float proj[16],
                             const float cam_pos[3]) {
    clear_error();
    DS_REQUIRE(r && view && proj && cam_pos, "NULL argument");
    DS_GUARD_BEGIN
    r->impl.render(view, proj, cam_pos);
    r->rendered = true;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_autotune(ds_renderer *r, int reps, int *fwd_pix_per_thread,
                               float *ms_per_frame) {
    clear_error();
    DS_REQUIRE(r, "NULL renderer");
    DS_REQUIRE(r->rendered, "ds_renderer_autotune needs one ds_renderer_render first");
    DS_GUARD_BEGIN
    float ms = 0.f;
    const int best = r->impl.autotune(reps, &ms);
    if (fwd_pix_per_thread) *fwd_pix_per_thread = best;
    if (ms_per_frame) *ms_per_frame = ms;
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_download_image(ds_renderer *r, float *host_rgb) {
    clear_error();
    DS_REQUIRE(r && host_rgb, "NULL argument");
    DS_TRY_HIP(hipStreamSynchronize(r->impl.getStream()));
    const size_t bytes = (size_t)r->impl.width() * r->impl.height() * 3 * sizeof(float);
    DS_TRY_HIP(hipMemcpy(host_rgb, r->impl.getOutput(), bytes, hipMemcpyDeviceToHost));
    return ds::raster_take_error(r->impl.raster(), r->impl.getStream());
}

ds_status ds_renderer_views(ds_renderer *r, ds_splat2d *splat2d) {
    clear_error();
    DS_REQUIRE(r && splat2d, "NULL argument");
    DS_GUARD_BEGIN
    *splat2d = r->impl.splat2dView();
    return DS_OK;
    DS_GUARD_END
}

ds_status ds_renderer_get_raster(ds_renderer *r, ds_raster **out) {
    clear_error();
    DS_REQUIRE(r && out, "NULL argument");
    *out = r->impl.raster();
    return DS_OK;
}

ds_status ds_renderer_get_stream(ds_renderer *r, ds_stream_t *s) {
    clear_error();
    DS_REQUIRE(r && s, "NULL argument");
    *s = r->impl.getStream();
    return DS_OK;
}

ds_status ds_renderer_synchronize(ds_renderer *r) {
    clear_error();
    DS_REQUIRE(r, "NULL renderer");
    DS_TRY_HIP(hipStreamSynchronize(r->impl.getStream()));
    return ds::raster_take_error(r->impl.raster(), r->impl.getStream());
}

}  // extern "C"
