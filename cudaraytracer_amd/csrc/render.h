// Host interface of the GPU scene and the wavefront renderer.
#pragma once

#include <stdint.h>
#include <vector>

#include "device.h"
#include "scene.h"
#include "svgf.h"

namespace rt {

enum Stage : int {
    kStageRaygen = 0,
    kStageExtend = 1,
    kStageShade = 2,
    kStageShadow = 3,
    kStageResolve = 4,
    kStageGBuffer = 5,
    kStageTemporal = 6,
    kStageVariance = 7,
    kStageAtrous = 8,
    kStageSort = 9,
    kNumStages = 10
};

// Buffers of the last real-time frame, for read-back.
enum FrameBuffer : int {
    kFrameDenoised = 0,          // rgba float: filtered direct + indirect, alpha 1
    kFrameDirect = 1,            // rgba float: filtered direct, .w variance
    kFrameIndirect = 2,
    kFrameHistoryDirect = 3,     // temporal output (the next frame's history), .w variance
    kFrameHistoryIndirect = 4,
    kFrameMomentsDirect = 5,     // float2 (mean luminance, mean squared luminance)
    kFrameMomentsIndirect = 6,
    kFrameHistoryLength = 7,     // int32
    kFrameVarianceDirect = 8,    // variance-pass output, .w variance
    kFrameVarianceIndirect = 9,
    kFrameGPosTri = 10,          // float4
    kFrameGNrmMat = 11,          // uint16 x4
    kFrameGBaryMesh = 12,        // uint16 x4
    kFrameGMotionDepth = 13,     // float4
    kNumFrameBuffers = 14
};
size_t frame_buffer_bytes(int which, int width, int height);

struct RenderSettings {
    int max_bounces = 5;
    float rr = 0.8f;
    int rr_start_depth = 0;
    uint32_t seed = 1234u;
    int batch_spp = 0;               // samples per pass (0: as many as the queue budget allows, up to 64)
    size_t queue_budget_bytes = 0;   // 0: 16 GiB
    int flush_survivors = 0;         // 1: paths alive after the last bounce keep their radiance
    int env_alias = 0;               // 1: environment sampled with alias tables (not the reference's CDF search)
    int use_graphs = 1;              // replay passes as captured HIP graphs
    int ray_sort = 0;                // 1: reorder extension rays by direction octant and origin cell
    int shadow_any_hit = 0;          // shadow test: 0 the reference's, 1 early exit (same results), 2 segment test
    int profile = 0;                 // per-stage HIP event timing
};

struct RenderStats {
    uint64_t ext_rays, shadow_rays, ext_nodes, ext_tris, shadow_nodes, shadow_tris, probe_rays;
    double stage_ms[kNumStages];
    int stage_launches[kNumStages];
    int spp_done, batch_spp;
    uint64_t miss_paths, shadow_lit;  // counting renders only
    uint64_t class_paths[kNumClasses]; // per shading class, counting renders only
};

struct EnvTables {
    std::vector<float> row_cdf, marg_cdf;
    // Alias tables: (threshold, alias index bits) per entry.
    std::vector<float> alias_row;   // 2 * env_w * env_h
    std::vector<float> alias_marg;  // 2 * env_h
};
// Walker/Vose alias table of n weights: threshold and alias per entry.
void build_alias_table(const double* w, int n, float* prob, int32_t* alias);
void build_env_tables(const EnvMap& env, EnvTables& t);

// Scene arrays resident in HBM.
class GpuScene {
public:
    explicit GpuScene(const HostScene& sc);
    ~GpuScene();
    GpuScene(const GpuScene&) = delete;
    GpuScene& operator=(const GpuScene&) = delete;
    const DevScene& dev() const;
    int max_stack() const;  // traversal stack entries the tree needs (depth + 1)
    bool has_class(int shading_class) const;
    size_t device_bytes() const;
    // Device side of a refit, after set_scene_geometry has updated `host`:
    // rewrites the triangle records (shading normals only when `nrm` is not
    // null), the light tables and every node box in place, then sets the
    // scene bounds of `host` and of dev() to the new root box.  The host
    // tree's other boxes are left stale; read_bvh brings them back.  Waits
    // for all work on the device first.  phase_ms: uploads, kernels.
    void refit(HostScene& host, const float* nrm, double phase_ms[2]);
    // Copy the device node boxes into b.packed and b.nodes.
    void read_bvh(BvhBuild& b) const;
    struct Impl;

private:
    Impl* impl_;
};

class Renderer {
public:
    Renderer(const GpuScene& scene, const HostScene& host, const RenderSettings& s);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    // Enqueue `spp` more samples per pixel on the renderer's stream.
    void render(int spp, bool count_traversal = false);
    // Camera of the following passes and frames; clears the accumulation
    // (a camera move resets the frame buffer: reference Src/Utils/ui.cpp:313-348).
    void set_camera(const CameraRec& c);
    const CameraRec& camera() const;
    // One real-time frame (reference: renderer.cuh:1141-1262 with denoising on):
    // G-buffer against the previous frame's camera, `spp` fresh samples
    // (sample indices frame * spp ...), temporal blend, variance estimate and
    // the a-trous iterations (skipped on the first frame after a history reset).
    void render_frame(int spp, const SvgfParams& p);
    // Forget the temporal history (hardResetDenoiseHistory, renderer.cuh:1125-1139).
    void reset_history();
    int frames() const;
    // Copy one of the last frame's buffers (FrameBuffer) into host memory of
    // frame_buffer_bytes(which, width, height) bytes.
    void read_frame(int which, void* out);
    void synchronize();
    void reset();
    void reset_stats();
    void stats(RenderStats& out);
    // which: 0 direct sum, 1 indirect sum, 2 (direct + indirect) / spp.  RGBA, row 0 = bottom.
    void read_accum(int which, float* out);
    int width() const;
    int height() const;
    int batch() const;
    int spp_done() const;
    struct Impl;

private:
    Impl* impl_;
};

// ------------------------------------------------------------------ pool
// Many render jobs over resident scenes, served by one wavefront engine: every
// pass carries the next work of as many jobs as its path capacity holds (all
// remaining samples of accumulate jobs that fit, one frame of each real-time
// job), so uneven jobs share each persistent stage launch.
struct PoolSettings {
    size_t queue_budget_bytes = 0;  // path queues; 0: 8 GiB
    size_t hbm_budget_bytes = 0;    // resident scenes + job state; 0: 64 GiB
    int profile = 0;
    int pipeline = 0;               // 1: two cohorts, one traced while the other is filtered
    int max_realtime = 0;           // serve's continuous batching admits at most this many real-time
                                    // jobs at once (0: as many as the HBM budget allows)
};

struct JobDesc {
    int scene = 0;                  // slot from add_scene
    RenderSettings set;
    int spp = 1;                    // accumulate: total samples; real-time: samples per frame
    int frames = 0;                 // 0: accumulate job; > 0: real-time job with SVGF
    SvgfParams svgf;
    std::vector<CameraRec> cams;    // real-time: one per frame; accumulate: at most one
};

struct JobStatus {
    int done = 0;
    int spp_done = 0;
    int frames_done = 0;
    double submit_s = 0, start_s = -1, done_s = -1;   // seconds on the pool's clock
    std::vector<double> frame_done_s;
};

class Pool {
public:
    explicit Pool(const PoolSettings& s);
    ~Pool();
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;
    // The scene objects must outlive the pool.
    int add_scene(const HostScene& host, const GpuScene& gpu);
    // Validate a job; returns the device bytes of its state.
    size_t check(const JobDesc& d) const;
    // Job id, or -1 when its state does not fit the HBM budget now.
    int submit(const JobDesc& d);
    const HostScene& scene_host(int slot) const;
    const GpuScene& scene_gpu(int slot) const;
    // One tick over the active jobs (synchronous); returns the number of jobs served.
    int step();
    int max_realtime() const;
    bool idle() const;
    const JobStatus& status(int job) const;
    void job_size(int job, int& width, int& height) const;
    void read_accum(int job, int which, float* out);
    void read_frame(int job, int which, void* out);
    void release(int job);
    double now() const;  // seconds since the pool was created
    size_t capacity() const;
    size_t bytes_in_use() const;
    void stats(RenderStats& out);
    void reset_stats();
    // Launch shape and pass size, chosen by autotune() or set directly.
    struct Tuning {
        float grid_scale = 1.f;   // persistent stage grids = occupancy x CUs x scale
        size_t pass_paths = 0;    // paths per pass (0: the engine's capacity)
        int tile_shift = 4;       // frame kernels on 2^shift x (256 >> shift) tiles
        int ray_sort = 0;         // 1: reorder extension rays before extend (38 B per path beyond the queue budget)
    };
    void set_tuning(const Tuning& t);
    Tuning tuning() const;
    // Time a sample of the expected jobs under candidate tunings (coordinate
    // descent: grid scale, then pass size, then tile shape, then the ray sort; each candidate
    // run twice, best time kept) and keep the fastest.  The pool must be idle.
    struct TuneReport {
        Tuning best;
        double baseline_ms = 0, best_ms = 0;
        int runs = 0;
    };
    TuneReport autotune(const std::vector<JobDesc>& sample);
    struct Impl;

private:
    int step_serial();
    int step_pipelined();
    Impl* impl_;
};

// Arrival-driven serving on a pool's scenes: each request is submitted at its
// arrival time (or later, when its state fits the HBM budget).  `batched`
// runs the pool's continuous batching; otherwise the plain loop renders the
// requests one at a time, in arrival order, each with its own fresh
// single-scene Renderer.  Frame latency = completion - (arrival for the first
// frame, completion of the previous frame otherwise); job latency =
// completion of the last frame - arrival.
struct ServeRequest {
    double arrival_s = 0;
    JobDesc job;
    // Optional, width * height * 4 floats: receives the job's result when it completes -- an
    // accumulate job's (direct + indirect) / spp, a real-time job's last denoised frame.
    float* result = nullptr;
};
// How serve() runs the stream: plain = one job at a time with a fresh
// single-scene renderer each; batched = the pool's continuous batching;
// plain_resident = one job at a time on a single-scene renderer kept per scene
// (history and accumulation reset per job).
enum ServeMode : int { kServePlain = 0, kServeBatched = 1, kServePlainResident = 2 };
struct ServeReport {
    int jobs = 0, frames = 0;
    double makespan_s = 0;                        // first arrival to last completion
    std::vector<std::vector<double>> frame_ms;    // per request, per frame
    std::vector<double> job_ms;                   // per request
};
ServeReport serve(Pool& pool, const std::vector<ServeRequest>& reqs, int mode);

// Closest hit for a batch of rays laid out as (o.xyz, tmin, d.xyz, tmax):
// out_hits = (t, b1, b2, prim bits) per ray, out_counts = (nodes, tris) per ray.
void trace_batch(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts);
// trace_batch without allocating device memory: rays, hits and counts are
// staged in a fixed device array of kNoAllocMaxRays rays (56 bytes each) on
// the current device.  Throws std::invalid_argument for n > kNoAllocMaxRays.
// Concurrent calls on one device share that array and race.
constexpr int kNoAllocMaxRays = 1 << 18;
constexpr int kMaxDevices = 64;
void trace_batch_noalloc(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts);
// Albedo of diffuse texture `tex` at n (u, v) pairs, on the GPU; 3 floats per pair.
void texture_lookup(const GpuScene& scene, const HostScene& host, int tex, const float* uv, int n, float* rgb_out);
void math_selftest(int op, const float* a, const float* b, float* out, int n);

}  // namespace rt
