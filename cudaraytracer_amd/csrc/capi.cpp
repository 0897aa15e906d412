// extern "C" boundary: opaque handles, return codes, thread-local last error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/rt.h"
#include "json.h"
#include "render.h"
#include "scene.h"

static_assert(RT_NUM_STAGES == rt::kNumStages, "stage ids of rt.h and the engine differ");
static_assert(RT_NUM_CLASSES == rt::kNumClasses, "shading classes of rt.h and the engine differ");
static_assert(RT_NOALLOC_MAX_RAYS == rt::kNoAllocMaxRays, "rt_trace_rays_noalloc limits of rt.h and the engine differ");

struct rt_scene {
    rt::HostScene host;
    std::unique_ptr<rt::GpuScene> gpu;  // created on first GPU use
    bool host_bvh_stale = false;        // a device refit left the host node boxes behind
    double refit_ms = 0.0;
    double refit_phase_ms[3] = {0, 0, 0};
    bool refit_on_gpu = false;
};

struct rt_renderer {
    rt_scene* scene = nullptr;
    std::unique_ptr<rt::Renderer> r;
};

struct rt_pool {
    std::unique_ptr<rt::Pool> p;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guard(F&& f) {
    try {
        g_err.clear();
        return f();
    } catch (const std::bad_alloc&) {
        return fail(RT_ERR_OUT_OF_MEMORY, "out of host memory");
    } catch (const std::invalid_argument& e) {
        return fail(RT_ERR_INVALID_ARG, e.what());
    } catch (const rt::JsonError& e) {
        return fail(RT_ERR_SCENE, e.what());
    } catch (const std::runtime_error& e) {
        std::string m = e.what();
        if (m.rfind("HIP error", 0) == 0) {
            if (m.find("out of memory") != std::string::npos) return fail(RT_ERR_OUT_OF_MEMORY, m);
            return fail(RT_ERR_DEVICE, m);
        }
        if (m.rfind("cannot open", 0) == 0) return fail(RT_ERR_IO, m);
        return fail(RT_ERR_SCENE, m);
    } catch (const std::exception& e) {
        return fail(RT_ERR_INTERNAL, e.what());
    }
}

void ensure_gpu(rt_scene* s) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw std::runtime_error("HIP error: no GPU device available");
    if (!s->gpu) s->gpu.reset(new rt::GpuScene(s->host));
}

// Render settings from the caller's overrides (negative = the scene's value).
rt::RenderSettings settings_from(const rt_render_params* p, const rt::HostScene& h) {
    rt_render_params d;
    rt_render_params_default(&d);
    if (!p) p = &d;
    rt::RenderSettings set;
    set.max_bounces = p->max_bounces >= 0 ? p->max_bounces : h.max_bounces;
    set.rr = p->rr >= 0.f ? p->rr : h.rr;
    set.rr_start_depth = p->rr_start_depth >= 0 ? p->rr_start_depth : h.rr_start_depth;
    set.seed = p->seed >= 0 ? (uint32_t)p->seed : h.seed;
    if (p->batch_spp < 0) throw std::invalid_argument("batch_spp must be >= 0");
    set.batch_spp = p->batch_spp;
    set.queue_budget_bytes = p->queue_budget_bytes > 0 ? (size_t)p->queue_budget_bytes : 0;
    set.flush_survivors = p->flush_survivors;
    set.profile = p->profile;
    if (p->env_sampling != 0 && p->env_sampling != 1) throw std::invalid_argument("env_sampling must be 0 or 1");
    set.env_alias = p->env_sampling;
    set.use_graphs = p->disable_graphs ? 0 : 1;
    if (p->ray_sort != 0 && p->ray_sort != 1) throw std::invalid_argument("ray_sort must be 0 or 1");
    set.ray_sort = p->ray_sort;
    if (p->shadow_any_hit < 0 || p->shadow_any_hit > 2) throw std::invalid_argument("shadow_any_hit must be 0, 1 or 2");
    set.shadow_any_hit = p->shadow_any_hit;
    if (set.max_bounces > 255) throw std::invalid_argument("max_bounces must be <= 255");
    return set;
}

// Orbit camera on top of a base camera (resolution, fov).
rt::CameraRec camera_from(const rt_camera* cam, const rt::CameraRec& base) {
    float vals[6] = {cam->target[0], cam->target[1], cam->target[2], cam->distance, cam->theta_deg, cam->phi_deg};
    for (float v : vals)
        if (!std::isfinite(v)) throw std::invalid_argument("camera parameters must be finite");
    if (!(cam->distance > 0.f)) throw std::invalid_argument("camera distance must be positive");
    rt::CameraRec c = base;
    c.target = rt::v3(cam->target[0], cam->target[1], cam->target[2]);
    c.distance = cam->distance;
    c.theta = cam->theta_deg;
    c.phi = cam->phi_deg;
    rt::compute_camera(c);
    return c;
}

rt::SvgfParams svgf_from(const rt_denoise_params* p) {
    rt::SvgfParams sp;
    if (p) {
        if (p->max_history < 1) throw std::invalid_argument("max_history must be >= 1");
        sp.max_history = p->max_history;
        sp.sigma_light = p->sigma_light;
        sp.sigma_normal = p->sigma_normal;
        sp.sigma_depth = p->sigma_depth;
        sp.reproj_depth = p->reprojection_depth;
        sp.atrous_iters = p->atrous_iterations;
    }
    return sp;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)std::ceil(q * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];  // nearest rank
}
}  // namespace

extern "C" {

void rt_render_params_default(rt_render_params* p) {
    if (!p) return;
    p->max_bounces = -1;
    p->rr = -1.f;
    p->rr_start_depth = -1;
    p->seed = -1;
    p->batch_spp = 0;
    p->queue_budget_bytes = 0;
    p->flush_survivors = 0;
    p->profile = 0;
    p->env_sampling = 0;
    p->disable_graphs = 0;
    p->ray_sort = 0;
    p->shadow_any_hit = 0;
}

void rt_denoise_params_default(rt_denoise_params* p) {
    if (!p) return;
    rt::SvgfParams d;
    p->max_history = d.max_history;
    p->sigma_light = d.sigma_light;
    p->sigma_normal = d.sigma_normal;
    p->sigma_depth = d.sigma_depth;
    p->reprojection_depth = d.reproj_depth;
    p->atrous_iterations = d.atrous_iters;
}

const char* rt_last_error(void) { return g_err.c_str(); }

int rt_scene_load(const char* json_path, const rt_load_opts* opts, rt_scene** out) {
    if (!json_path || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (opts && (opts->bvh_builder < 0 || opts->bvh_builder > 2))
        return fail(RT_ERR_INVALID_ARG, "rt_load_opts.bvh_builder must be 0, 1 or 2");
    *out = nullptr;
    return guard([&]() {
        std::unique_ptr<rt_scene> s(new rt_scene);
        rt::load_scene(json_path, s->host, opts ? opts->bvh_builder : rt::kBvhAuto);
        *out = s.release();
        return RT_OK;
    });
}

int rt_scene_info_get(const rt_scene* s, rt_scene_info* o) {
    if (!s || !o) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::HostScene& h = s->host;
    std::memset(o, 0, sizeof(*o));
    o->num_triangles = h.num_tris();
    o->num_meshes = (int)h.meshes.size();
    o->num_materials = (int)h.materials.size();
    o->num_lights = (int)h.lights.size();
    o->num_bvh_nodes = (int)h.bvh.nodes.size();
    o->bvh_max_depth = h.bvh.max_depth;
    o->bvh_fallback_splits = h.bvh.fallback_splits;
    o->width = h.camera.width;
    o->height = h.camera.height;
    o->spp = h.spp;
    o->max_bounces = h.max_bounces;
    o->rr = h.rr;
    o->denoise = h.denoise ? 1 : 0;
    o->seed = h.seed;
    o->rr_start_depth = h.rr_start_depth;
    o->num_textures = (int)h.textures.size();
    o->bvh_build_ms = (float)h.bvh_build_ms;
    o->bvh_on_gpu = h.bvh_on_gpu ? 1 : 0;
    for (int k = 0; k < 4; ++k) o->bvh_gpu_phase_ms[k] = (float)h.bvh.gpu_phase_ms[k];
    for (int k = 0; k < 3; ++k) o->load_phase_ms[k] = (float)h.load_phase_ms[k];
    o->refit_ms = (float)s->refit_ms;
    o->refit_on_gpu = s->refit_on_gpu ? 1 : 0;
    for (int k = 0; k < 3; ++k) o->refit_phase_ms[k] = (float)s->refit_phase_ms[k];
    const rt::Box& b = h.bvh.bounds;
    float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
    std::memcpy(o->bounds_lo, lo, sizeof(lo));
    std::memcpy(o->bounds_hi, hi, sizeof(hi));
    const rt::CameraRec& c = h.camera;
    float cp[3] = {c.position.x, c.position.y, c.position.z};
    float ld[3] = {c.left_down.x, c.left_down.y, c.left_down.z};
    float du[3] = {c.du.x, c.du.y, c.du.z};
    float dv[3] = {c.dv.x, c.dv.y, c.dv.z};
    std::memcpy(o->camera_position, cp, sizeof(cp));
    std::memcpy(o->camera_left_down, ld, sizeof(ld));
    std::memcpy(o->camera_du, du, sizeof(du));
    std::memcpy(o->camera_dv, dv, sizeof(dv));
    return RT_OK;
}

void rt_scene_destroy(rt_scene* s) { delete s; }

int rt_scene_refit(rt_scene* s, const float* pos, const float* nrm, int64_t num_triangles) {
    if (!s || !pos) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (num_triangles != (int64_t)s->host.num_tris())
        return fail(RT_ERR_INVALID_ARG, "rt_scene_refit: " + std::to_string(num_triangles) +
                                            " triangles given, the scene has " + std::to_string(s->host.num_tris()));
    return guard([&]() {
        auto t0 = std::chrono::steady_clock::now();
        rt::set_scene_geometry(s->host, pos, nrm);
        rt::refit_media(s->host);
        auto t1 = std::chrono::steady_clock::now();
        s->refit_phase_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (s->gpu) {
            s->gpu->refit(s->host, nrm, s->refit_phase_ms + 1);
            s->host_bvh_stale = true;
        } else {
            rt::refit_bvh(s->host.pos.data(), s->host.num_tris(), s->host.bvh);
            s->refit_phase_ms[1] = 0.0;
            s->refit_phase_ms[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        }
        s->refit_on_gpu = (bool)s->gpu;
        s->refit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return RT_OK;
    });
}

int rt_renderer_create(rt_scene* s, const rt_render_params* p, rt_renderer** out) {
    if (!s || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    return guard([&]() {
        rt::RenderSettings set = settings_from(p, s->host);
        ensure_gpu(s);
        std::unique_ptr<rt_renderer> r(new rt_renderer);
        r->scene = s;
        r->r.reset(new rt::Renderer(*s->gpu, s->host, set));
        *out = r.release();
        return RT_OK;
    });
}

void rt_renderer_destroy(rt_renderer* r) { delete r; }

int rt_render(rt_renderer* r, int spp) {
    if (!r || spp < 0) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        r->r->render(spp, false);
        return RT_OK;
    });
}

int rt_render_counted(rt_renderer* r, int spp) {
    if (!r || spp < 0) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        r->r->render(spp, true);
        return RT_OK;
    });
}

int rt_synchronize(rt_renderer* r) {
    if (!r) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        r->r->synchronize();
        return RT_OK;
    });
}

int rt_reset(rt_renderer* r) {
    if (!r) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        r->r->reset();
        return RT_OK;
    });
}

int rt_reset_stats(rt_renderer* r) {
    if (!r) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        r->r->reset_stats();
        return RT_OK;
    });
}

int rt_read_accum(rt_renderer* r, int which, float* rgba) {
    if (!r || !rgba || which < 0 || which > 2) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        r->r->read_accum(which, rgba);
        return RT_OK;
    });
}

// GatherKernel colour (renderer.cuh:969-988): gamma 1/2.2, clamp, truncate to
// 8 bits, alpha 255, rows flipped to top first.  `img` is RGBA float, row 0 = bottom.
static void gather_rgba8(const float* img, int w, int h, uint8_t* out) {
    for (int j = 0; j < h; ++j) {
        int dst = h - 1 - j;  // top row first
        for (int i = 0; i < w; ++i) {
            const float* c = &img[((size_t)j * w + i) * 4];
            uint8_t* o = &out[((size_t)dst * w + i) * 4];
            for (int k = 0; k < 3; ++k) {
                float v = std::pow(c[k], 1.0f / 2.2f) * 255.0f;
                v = std::fmin(std::fmax(v, 0.0f), 255.0f);
                o[k] = (uint8_t)v;
            }
            o[3] = 255;
        }
    }
}

int rt_read_rgba8(rt_renderer* r, uint8_t* out) {
    if (!r || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        int w = r->r->width(), h = r->r->height();
        std::vector<float> img((size_t)w * h * 4);
        r->r->read_accum(RT_ACCUM_MEAN, img.data());
        gather_rgba8(img.data(), w, h, out);
        return RT_OK;
    });
}

int rt_read_denoised(rt_renderer* r, float* rgba_out) {
    if (!r || !rgba_out) return fail(RT_ERR_INVALID_ARG, "null argument");
    return rt_read_frame(r, RT_FRAME_DENOISED, rgba_out, (size_t)r->r->width() * r->r->height() * 16);
}

int rt_read_gbuffer(rt_renderer* r, int channel, void* out) {
    if (!r || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (channel < 0 || channel > 3) return fail(RT_ERR_INVALID_ARG, "G-buffer channel must be 0..3");
    const int which = RT_FRAME_G_POSITION_TRI + channel;
    return rt_read_frame(r, which, out, rt::frame_buffer_bytes(which, r->r->width(), r->r->height()));
}

int rt_read_gather(rt_renderer* r, int mode, uint8_t* out) {
    if (!r || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (mode < RT_GATHER_COLOR || mode > RT_GATHER_DENOISED) return fail(RT_ERR_INVALID_ARG, "unknown gather mode");
    if (mode == RT_GATHER_COLOR) return rt_read_rgba8(r, out);
    return guard([&]() {
        const int w = r->r->width(), h = r->r->height();
        if (r->r->frames() == 0) throw std::invalid_argument("no real-time frame has been rendered");
        std::vector<float> f4((size_t)w * h * 4);
        if (mode == RT_GATHER_DENOISED) {
            r->r->read_frame(RT_FRAME_DENOISED, f4.data());
            gather_rgba8(f4.data(), w, h, out);
            return RT_OK;
        }
        // GatherKernel's render-target modes 1-4 (renderer.cuh:990-1029) read the
        // G-buffer row h-1-j into output row j: top row first, like the colour mode.
        std::vector<uint16_t> u4((size_t)w * h * 4);
        const rt::Box& b = r->scene->host.bvh.bounds;
        const rt::CameraRec& cam = r->r->camera();
        auto u8 = [](float v) { return (uint8_t)std::fmin(std::fmax(v, 0.0f), 255.0f); };
        if (mode == RT_GATHER_DEPTH) r->r->read_frame(RT_FRAME_G_MOTION_DEPTH, f4.data());
        else if (mode == RT_GATHER_POSITION) r->r->read_frame(RT_FRAME_G_POSITION_TRI, f4.data());
        else r->r->read_frame(mode == RT_GATHER_NORMAL ? RT_FRAME_G_NORMAL_MAT : RT_FRAME_G_BARY_MESH, u4.data());
        // Depth range: near 0.1, far = scene diagonal + camera distance to the scene centre.
        rt::vec3 ext = b.hi - b.lo, centre = (b.lo + b.hi) * 0.5f, off = cam.position - centre;
        const float far_plane = rt::length(ext) + rt::length(off);  // Eigen norm order, as everywhere
        for (int j = 0; j < h; ++j) {
            const size_t src_row = (size_t)(h - 1 - j) * w;
            for (int i = 0; i < w; ++i) {
                const size_t s = src_row + i;
                uint8_t* o = out + 4 * ((size_t)j * w + i);
                o[3] = 255;
                if (mode == RT_GATHER_DEPTH) {
                    float d = (f4[4 * s + 2] - 0.1f) / (far_plane - 0.1f);
                    o[0] = o[1] = o[2] = u8(d * 255.0f);
                } else if (mode == RT_GATHER_NORMAL) {
                    for (int k = 0; k < 3; ++k) o[k] = u8(((float)u4[4 * s + k] / 65535.0f) * 255.0f);
                } else if (mode == RT_GATHER_ID) {
                    const unsigned id = u4[4 * s + 3];
                    o[0] = (uint8_t)((id * 37) % 256);
                    o[1] = (uint8_t)((id * 57) % 256);
                    o[2] = (uint8_t)((id * 97) % 256);
                } else {
                    const float lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
                    for (int k = 0; k < 3; ++k) o[k] = u8((f4[4 * s + k] - lo[k]) / (hi[k] - lo[k]) * 255.0f);
                }
            }
        }
        return RT_OK;
    });
}

int rt_save_image(rt_renderer* r, int which, const char* path) {
    if (!r || !path) return fail(RT_ERR_INVALID_ARG, "null argument");
    if (which < RT_IMAGE_PNG_COLOR || which > RT_IMAGE_PFM_DENOISED) return fail(RT_ERR_INVALID_ARG, "unknown image kind");
    return guard([&]() {
        int w = r->r->width(), h = r->r->height();
        const bool denoised = which == RT_IMAGE_PNG_DENOISED || which == RT_IMAGE_PFM_DENOISED;
        if (denoised && r->r->frames() == 0) throw std::invalid_argument("no real-time frame has been rendered");
        std::vector<float> img((size_t)w * h * 4);
        if (denoised) r->r->read_frame(RT_FRAME_DENOISED, img.data());
        else r->r->read_accum(RT_ACCUM_MEAN, img.data());
        if (which == RT_IMAGE_PNG_COLOR || which == RT_IMAGE_PNG_DENOISED) {
            std::vector<uint8_t> px((size_t)w * h * 4);
            gather_rgba8(img.data(), w, h, px.data());
            rt::write_png_rgba8(path, w, h, px.data());
        } else {
            rt::write_pfm_rgb(path, w, h, img.data());
        }
        return RT_OK;
    });
}

int rt_stats(rt_renderer* r, rt_stats_out* o) {
    if (!r || !o) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        rt::RenderStats s;
        r->r->stats(s);
        o->extension_rays = s.ext_rays;
        o->shadow_rays = s.shadow_rays;
        o->extension_node_visits = s.ext_nodes;
        o->extension_tri_tests = s.ext_tris;
        o->shadow_node_visits = s.shadow_nodes;
        o->shadow_tri_tests = s.shadow_tris;
        o->probe_rays = s.probe_rays;
        o->miss_paths = s.miss_paths;
        o->shadow_lit = s.shadow_lit;
        for (int k = 0; k < RT_NUM_CLASSES; ++k) o->class_paths[k] = s.class_paths[k];
        for (int k = 0; k < RT_NUM_STAGES; ++k) {
            o->stage_ms[k] = s.stage_ms[k];
            o->stage_launches[k] = s.stage_launches[k];
        }
        o->spp_done = s.spp_done;
        o->batch_spp = s.batch_spp;
        return RT_OK;
    });
}

int rt_get_camera(rt_renderer* r, rt_camera* out) {
    if (!r || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::CameraRec& c = r->r->camera();
    out->target[0] = c.target.x;
    out->target[1] = c.target.y;
    out->target[2] = c.target.z;
    out->distance = c.distance;
    out->theta_deg = c.theta;
    out->phi_deg = c.phi;
    return RT_OK;
}

int rt_set_camera(rt_renderer* r, const rt_camera* cam) {
    if (!r || !cam) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        r->r->set_camera(camera_from(cam, r->r->camera()));
        return RT_OK;
    });
}

int rt_render_frame(rt_renderer* r, int spp, const rt_denoise_params* p) {
    if (!r || spp < 1) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        r->r->render_frame(spp, svgf_from(p));
        return RT_OK;
    });
}

int rt_reset_history(rt_renderer* r) {
    if (!r) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        r->r->reset_history();
        return RT_OK;
    });
}

int rt_read_frame(rt_renderer* r, int which, void* out, size_t bytes) {
    if (!r || !out || which < 0 || which >= rt::kNumFrameBuffers) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    size_t need = rt::frame_buffer_bytes(which, r->r->width(), r->r->height());
    if (bytes != need)
        return fail(RT_ERR_INVALID_ARG, "buffer size " + std::to_string(bytes) + " != " + std::to_string(need));
    return guard([&]() {
        r->r->read_frame(which, out);
        return RT_OK;
    });
}

// ------------------------------------------------------------------ pools
namespace {
rt::JobDesc job_from(rt_pool* p, const rt_job_desc* d) {
    const rt::HostScene& h = p->p->scene_host(d->scene);
    rt::JobDesc j;
    j.scene = d->scene;
    j.set = settings_from(d->render, h);
    j.spp = d->spp;
    j.frames = d->frames;
    j.svgf = svgf_from(d->denoise);
    int ncam = d->frames > 0 ? d->frames : (d->cameras ? 1 : 0);
    if (d->frames > 0 && !d->cameras) throw std::invalid_argument("a real-time job needs one camera per frame");
    for (int k = 0; k < ncam; ++k) j.cams.push_back(camera_from(&d->cameras[k], h.camera));
    return j;
}
}  // namespace

int rt_pool_create(const rt_pool_params* p, rt_pool** out) {
    if (!out) return fail(RT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (p && (p->queue_budget_bytes < 0 || p->hbm_budget_bytes < 0)) return fail(RT_ERR_INVALID_ARG, "negative budget");
    return guard([&]() {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw std::runtime_error("HIP error: no GPU device available");
        rt::PoolSettings s;
        if (p) {
            s.queue_budget_bytes = (size_t)p->queue_budget_bytes;
            s.hbm_budget_bytes = (size_t)p->hbm_budget_bytes;
            s.profile = p->profile;
            s.pipeline = p->cohort_ticks ? 1 : 0;
            if (p->max_realtime_jobs < 0) throw std::invalid_argument("max_realtime_jobs must be >= 0");
            s.max_realtime = p->max_realtime_jobs;
        }
        std::unique_ptr<rt_pool> pool(new rt_pool);
        pool->p.reset(new rt::Pool(s));
        *out = pool.release();
        return RT_OK;
    });
}

void rt_pool_destroy(rt_pool* p) { delete p; }

int rt_pool_add_scene(rt_pool* p, rt_scene* s, int* slot_out) {
    if (!p || !s || !slot_out) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        ensure_gpu(s);
        *slot_out = p->p->add_scene(s->host, *s->gpu);
        return RT_OK;
    });
}

int rt_pool_submit(rt_pool* p, const rt_job_desc* d, int* job_out) {
    if (!p || !d || !job_out) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        int id = p->p->submit(job_from(p, d));
        if (id < 0) return fail(RT_ERR_BUSY, "the job's state does not fit the pool's HBM budget now");
        *job_out = id;
        return RT_OK;
    });
}

int rt_pool_step(rt_pool* p, int* served_out) {
    if (!p) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        int n = p->p->step();
        if (served_out) *served_out = n;
        return RT_OK;
    });
}

int rt_pool_job_status(rt_pool* p, int job, rt_job_status* out) {
    if (!p || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        const rt::JobStatus& st = p->p->status(job);
        out->done = st.done;
        out->spp_done = st.spp_done;
        out->frames_done = st.frames_done;
        out->submit_s = st.submit_s;
        out->start_s = st.start_s;
        out->done_s = st.done_s;
        return RT_OK;
    });
}

int rt_pool_read_accum(rt_pool* p, int job, int which, float* rgba) {
    if (!p || !rgba || which < 0 || which > 2) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        p->p->read_accum(job, which, rgba);
        return RT_OK;
    });
}

int rt_pool_read_frame(rt_pool* p, int job, int which, void* out, size_t bytes) {
    if (!p || !out || which < 0 || which >= rt::kNumFrameBuffers) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        int w = 0, h = 0;
        p->p->job_size(job, w, h);
        size_t need = rt::frame_buffer_bytes(which, w, h);
        if (bytes != need)
            throw std::invalid_argument("buffer size " + std::to_string(bytes) + " != " + std::to_string(need));
        p->p->read_frame(job, which, out);
        return RT_OK;
    });
}

int rt_pool_release(rt_pool* p, int job) {
    if (!p) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        p->p->release(job);
        return RT_OK;
    });
}

int rt_pool_stats(rt_pool* p, rt_stats_out* o) {
    if (!p || !o) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        rt::RenderStats s;
        p->p->stats(s);
        std::memset(o, 0, sizeof(*o));
        o->extension_rays = s.ext_rays;
        o->shadow_rays = s.shadow_rays;
        o->extension_node_visits = s.ext_nodes;
        o->extension_tri_tests = s.ext_tris;
        o->shadow_node_visits = s.shadow_nodes;
        o->shadow_tri_tests = s.shadow_tris;
        o->probe_rays = s.probe_rays;
        o->miss_paths = s.miss_paths;
        o->shadow_lit = s.shadow_lit;
        for (int k = 0; k < RT_NUM_CLASSES; ++k) o->class_paths[k] = s.class_paths[k];
        for (int k = 0; k < RT_NUM_STAGES; ++k) {
            o->stage_ms[k] = s.stage_ms[k];
            o->stage_launches[k] = s.stage_launches[k];
        }
        return RT_OK;
    });
}

int rt_pool_reset_stats(rt_pool* p) {
    if (!p) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        p->p->reset_stats();
        return RT_OK;
    });
}

namespace {
int tile_shift_of(int width) {
    for (int s = 3; s <= 7; ++s)
        if ((1 << s) == width) return s;
    throw std::invalid_argument("tile_width must be 8, 16, 32, 64 or 128");
}
void tuning_out(const rt::Pool::Tuning& t, rt_pool_tuning* o) {
    o->grid_scale = t.grid_scale;
    o->pass_paths = (int64_t)t.pass_paths;
    o->tile_width = 1 << t.tile_shift;
    o->ray_sort = t.ray_sort;
}
}  // namespace

int rt_pool_set_tuning(rt_pool* p, const rt_pool_tuning* t) {
    if (!p || !t) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        if (t->pass_paths < 0) throw std::invalid_argument("pass_paths must be >= 0");
        rt::Pool::Tuning tu;
        tu.grid_scale = t->grid_scale;
        tu.pass_paths = (size_t)t->pass_paths;
        tu.tile_shift = tile_shift_of(t->tile_width);
        tu.ray_sort = t->ray_sort;
        p->p->set_tuning(tu);
        return RT_OK;
    });
}

int rt_pool_get_tuning(rt_pool* p, rt_pool_tuning* out) {
    if (!p || !out) return fail(RT_ERR_INVALID_ARG, "null argument");
    tuning_out(p->p->tuning(), out);
    return RT_OK;
}

int rt_pool_autotune(rt_pool* p, const rt_job_desc* sample, int n, rt_tune_report* out) {
    if (!p || !sample || n < 1 || !out) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        std::vector<rt::JobDesc> jobs;
        for (int k = 0; k < n; ++k) jobs.push_back(job_from(p, &sample[k]));
        rt::Pool::TuneReport r = p->p->autotune(jobs);
        tuning_out(r.best, &out->best);
        out->baseline_ms = r.baseline_ms;
        out->best_ms = r.best_ms;
        out->runs = r.runs;
        return RT_OK;
    });
}

int rt_pool_serve(rt_pool* p, const rt_serve_request* reqs, int n, int batched, rt_serve_report* out,
                  double* frame_latency_ms, double* job_latency_ms) {
    if (!p || (!reqs && n > 0) || n < 0 || !out) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    return guard([&]() {
        std::vector<rt::ServeRequest> rq((size_t)n);
        for (int k = 0; k < n; ++k) {
            if (!(reqs[k].arrival_s >= 0.0)) throw std::invalid_argument("arrival times must be >= 0");
            rq[k].arrival_s = reqs[k].arrival_s;
            rq[k].job = job_from(p, &reqs[k].job);
            rq[k].result = reqs[k].result_rgba;
        }
        rt::ServeReport rep = rt::serve(*p->p, rq, batched);
        std::vector<double> frames;
        for (const auto& f : rep.frame_ms) frames.insert(frames.end(), f.begin(), f.end());
        std::memset(out, 0, sizeof(*out));
        out->jobs = rep.jobs;
        out->frames = rep.frames;
        out->makespan_s = rep.makespan_s;
        out->jobs_per_s = rep.makespan_s > 0 ? rep.jobs / rep.makespan_s : 0.0;
        out->frame_p50_ms = percentile(frames, 0.50);
        out->frame_p99_ms = percentile(frames, 0.99);
        out->job_p50_ms = percentile(rep.job_ms, 0.50);
        out->job_p99_ms = percentile(rep.job_ms, 0.99);
        if (frame_latency_ms) std::copy(frames.begin(), frames.end(), frame_latency_ms);
        if (job_latency_ms) std::copy(rep.job_ms.begin(), rep.job_ms.end(), job_latency_ms);
        return RT_OK;
    });
}

int rt_scene_get_triangles(const rt_scene* s, float* pos, float* nrm, float* uv, int32_t* mesh) {
    if (!s) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::HostScene& h = s->host;
    if (pos) std::memcpy(pos, h.pos.data(), h.pos.size() * 4);
    if (nrm) std::memcpy(nrm, h.nrm.data(), h.nrm.size() * 4);
    if (uv) std::memcpy(uv, h.uv.data(), h.uv.size() * 4);
    if (mesh) std::memcpy(mesh, h.mesh.data(), h.mesh.size() * 4);
    return RT_OK;
}

int rt_scene_get_meshes(const rt_scene* s, int32_t* mesh4, float* area) {
    if (!s) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::HostScene& h = s->host;
    for (size_t m = 0; m < h.meshes.size(); ++m) {
        if (mesh4) {
            mesh4[4 * m] = h.meshes[m].material;
            mesh4[4 * m + 1] = h.meshes[m].light;
            mesh4[4 * m + 2] = h.meshes[m].start;
            mesh4[4 * m + 3] = h.meshes[m].count;
        }
        if (area) area[m] = h.meshes[m].area;
    }
    return RT_OK;
}

int rt_scene_get_bvh(const rt_scene* s, int from_packed, void* nodes_out, int32_t* order_out) {
    if (!s) return fail(RT_ERR_INVALID_ARG, "null argument");
    return guard([&]() {
        if (s->host_bvh_stale) {  // bring the device refit's boxes back to the host tree
            rt_scene* m = const_cast<rt_scene*>(s);
            m->gpu->read_bvh(m->host.bvh);
            m->host_bvh_stale = false;
        }
        const rt::BvhBuild& b = s->host.bvh;
        if (nodes_out) {
            if (from_packed) {
                std::vector<rt::LinearNode> un;
                rt::unpack_bvh(b.packed, b.root_ref, b.bounds, un);
                if (un.size() != b.nodes.size()) throw std::runtime_error("unpacked node count differs");
                std::memcpy(nodes_out, un.data(), un.size() * sizeof(rt::LinearNode));
            } else {
                std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(rt::LinearNode));
            }
        }
        if (order_out) std::memcpy(order_out, b.order.data(), b.order.size() * 4);
        return RT_OK;
    });
}

int rt_scene_get_texture(const rt_scene* s, int index, int* w, int* h, uint8_t* rgba) {
    if (!s || !w || !h) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::HostScene& hs = s->host;
    if (index < 0 || index >= (int)hs.textures.size()) return fail(RT_ERR_INVALID_ARG, "texture index out of range");
    const rt::LdrImage& t = hs.textures[index];
    *w = t.w;
    *h = t.h;
    if (rgba && !t.rgba.empty()) std::memcpy(rgba, t.rgba.data(), t.rgba.size());
    return RT_OK;
}

int rt_scene_get_materials(const rt_scene* s, int* n, void* out) {
    if (!s || !n) return fail(RT_ERR_INVALID_ARG, "null argument");
    const std::vector<rt::MaterialRec>& m = s->host.materials;
    *n = (int)m.size();
    if (out && !m.empty()) std::memcpy(out, m.data(), m.size() * sizeof(rt::MaterialRec));
    return RT_OK;
}

int rt_scene_get_medium(const rt_scene* s, int mesh, int* n, float* lo, float* inv, float* cells) {
    if (!s || !n) return fail(RT_ERR_INVALID_ARG, "null argument");
    *n = 0;
    for (const rt::MediumTree& t : s->host.media) {
        if (t.mesh != mesh) continue;
        *n = t.grid_n;
        if (lo) { lo[0] = t.grid_lo.x; lo[1] = t.grid_lo.y; lo[2] = t.grid_lo.z; }
        if (inv) { inv[0] = t.grid_inv.x; inv[1] = t.grid_inv.y; inv[2] = t.grid_inv.z; }
        if (cells) std::memcpy(cells, t.clearance.data(), t.clearance.size() * sizeof(float));
    }
    return RT_OK;
}

int rt_scene_get_envmap(const rt_scene* s, int* w, int* h, float* rgb) {
    if (!s || !w || !h) return fail(RT_ERR_INVALID_ARG, "null argument");
    const rt::EnvMap& e = s->host.env;
    *w = e.width;
    *h = e.height;
    if (rgb && !e.rgb.empty()) std::memcpy(rgb, e.rgb.data(), e.rgb.size() * sizeof(float));
    return RT_OK;
}

int rt_texture_lookup(rt_scene* s, int index, const float* uv, int n, float* rgb_out) {
    if (!s || !uv || !rgb_out || n < 0) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    if (n == 0) return RT_OK;
    return guard([&]() {
        ensure_gpu(s);
        rt::texture_lookup(*s->gpu, s->host, index, uv, n, rgb_out);
        return RT_OK;
    });
}

int rt_trace_rays(rt_scene* s, const float* rays, int n, float* hits, uint32_t* counts) {
    if (!s || !rays || !hits || n < 0) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    if (n == 0) return RT_OK;
    return guard([&]() {
        ensure_gpu(s);
        rt::trace_batch(*s->gpu, rays, n, hits, counts);
        return RT_OK;
    });
}

int rt_trace_rays_noalloc(rt_scene* s, const float* rays, int n, float* hits, uint32_t* counts) {
    if (!s || !rays || !hits || n < 0) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    if (n > RT_NOALLOC_MAX_RAYS)
        return fail(RT_ERR_INVALID_ARG, "rt_trace_rays_noalloc takes at most " + std::to_string(RT_NOALLOC_MAX_RAYS) +
                                            " rays per call, got " + std::to_string(n));
    if (!s->gpu)
        return fail(RT_ERR_INVALID_ARG,
                    "the scene is not resident on the device: create a renderer on it or call rt_trace_rays first");
    if (n == 0) return RT_OK;
    return guard([&]() {
        rt::trace_batch_noalloc(*s->gpu, rays, n, hits, counts);
        return RT_OK;
    });
}

int rt_device_math(int op, const float* a, const float* b, float* out, int n) {
    if (!a || !b || !out || n < 0 || op < 0 || op > 10) return fail(RT_ERR_INVALID_ARG, "invalid argument");
    if (n == 0) return RT_OK;
    return guard([&]() {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) throw std::runtime_error("HIP error: no GPU device available");
        rt::math_selftest(op, a, b, out, n);
        return RT_OK;
    });
}

}  // extern "C"
