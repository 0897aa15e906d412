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
    kNumStages = 9
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
    size_t queue_budget_bytes = 0;   // 0: 8 GiB
    int flush_survivors = 0;         // 1: paths alive after the last bounce keep their radiance
    int profile = 0;                 // per-stage HIP event timing
};

struct RenderStats {
    uint64_t ext_rays, shadow_rays, ext_nodes, ext_tris, shadow_nodes, shadow_tris, probe_rays;
    double stage_ms[kNumStages];
    int stage_launches[kNumStages];
    int spp_done, batch_spp;
};

struct EnvTables {
    std::vector<float> row_cdf, marg_cdf;
};
void build_env_tables(const EnvMap& env, EnvTables& t);

// Scene arrays resident in HBM.
class GpuScene {
public:
    explicit GpuScene(const HostScene& sc);
    ~GpuScene();
    GpuScene(const GpuScene&) = delete;
    GpuScene& operator=(const GpuScene&) = delete;
    const DevScene& dev() const;
    int stack() const;
    bool has_class(int shading_class) const;
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

// Closest hit for a batch of rays laid out as (o.xyz, tmin, d.xyz, tmax):
// out_hits = (t, b1, b2, prim bits) per ray, out_counts = (nodes, tris) per ray.
void trace_batch(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts);
void math_selftest(int op, const float* a, const float* b, float* out, int n);

}  // namespace rt
