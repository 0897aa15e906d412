/*
 * C ABI of the MI355X wavefront path tracer.
 *
 * The reference has no headless API: its only caller is main()
 * (Src/main.cu:18-45), which constructs Scene(path) (Src/Core/scene.cpp:50-57),
 * then Renderer(ui) (Src/Core/renderer.cuh:1094-1100) and runs RenderLoop()
 * (Src/Core/renderer.cuh:1106-1307); results live in the device accumulators
 * renderBuffers[].{direct,indirect}LightingBuffer (renderer.cuh:56-59).  The
 * entry points below replace that surface one to one:
 *
 *   rt_scene_load        <- Scene::Scene (scene.cpp:50-57): JSON + OBJ, SAH BVH
 *   rt_scene_refit       <- a new Scene for moved vertices (no refit in the reference)
 *   rt_renderer_create   <- Renderer::Renderer / CudaInit (renderer.cuh:1094-1100, 1309-1355)
 *   rt_render            <- RenderLoop frames without denoising (renderer.cuh:1144-1193),
 *                           one reference frame = one sample per pixel
 *   rt_read_accum        <- renderBuffers[cur].direct/indirectLightingBuffer read-back
 *   rt_read_rgba8        <- GatherKernel colour mode (renderer.cuh:956-993)
 *   rt_stats             <- DenoiseTimingStats / stage timers (ui.h:27-37), plus ray counts
 *   rt_set_camera        <- orbit-camera update (Src/Utils/ui.cpp:313-348 -> camera.h:58-84)
 *   rt_render_frame      <- RenderLoop frame with denoising on (renderer.cuh:1141-1262):
 *                           G-buffer (renderer.cuh:1446-1503), path trace, temporal,
 *                           variance and a-trous kernels (renderer.cuh:442-954)
 *   rt_read_frame        <- GatherKernel denoise branch (renderer.cuh:970-978) and the
 *                           G-buffer / history buffers the SVGF kernels read
 *   rt_reset_history     <- hardResetDenoiseHistory (renderer.cuh:1125-1139)
 *   rt_pool_*            <- no counterpart: the reference renders one scene per process
 *                           (main.cu:26-33).  A pool keeps many scenes resident and
 *                           runs many jobs through one wavefront engine (batching), and
 *                           rt_pool_serve replays an arrival stream against it (serving).
 *   rt_last_error        <- std::runtime_error / printed CUDA errors (scene.cpp:51-53, global.h:19-27)
 *
 * Conventions: functions return RT_OK (0) or a negative RT_ERR_* code and set
 * a thread-local message readable with rt_last_error().  Output buffers are
 * caller-owned host memory; device memory belongs to the library.  Images are
 * RGBA float32, width*height*4, row 0 = the BOTTOM row of the image (the
 * reference's accumulation order); rt_read_rgba8 flips to top-first like the
 * reference's display gather.  A handle may be used by one thread at a time;
 * distinct handles are independent and each renderer owns one HIP stream.
 */
#ifndef RT_AMD_H
#define RT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_OK 0
#define RT_ERR_INVALID_ARG (-1)
#define RT_ERR_IO (-2)
#define RT_ERR_SCENE (-3)
#define RT_ERR_DEVICE (-4)
#define RT_ERR_OUT_OF_MEMORY (-5)
#define RT_ERR_INTERNAL (-6)
#define RT_ERR_BUSY (-7) /* the pool's HBM budget cannot admit the job now */

#define RT_ACCUM_DIRECT 0
#define RT_ACCUM_INDIRECT 1
#define RT_ACCUM_MEAN 2 /* (direct + indirect) / spp */

#define RT_GATHER_COLOR 0     /* renderTargetMode 0, accumulated */
#define RT_GATHER_DEPTH 1
#define RT_GATHER_NORMAL 2
#define RT_GATHER_ID 3
#define RT_GATHER_POSITION 4
#define RT_GATHER_DENOISED 5  /* renderTargetMode 0 with denoising on */

#define RT_IMAGE_PNG_COLOR 0
#define RT_IMAGE_PNG_DENOISED 1
#define RT_IMAGE_PFM_MEAN 2
#define RT_IMAGE_PFM_DENOISED 3

#define RT_STAGE_RAYGEN 0
#define RT_STAGE_EXTEND 1
#define RT_STAGE_SHADE 2
#define RT_STAGE_SHADOW 3
#define RT_STAGE_RESOLVE 4
#define RT_STAGE_GBUFFER 5
#define RT_STAGE_TEMPORAL 6
#define RT_STAGE_VARIANCE 7
#define RT_STAGE_ATROUS 8
#define RT_STAGE_SORT 9 /* ray sort before extend (rt_render_params.ray_sort) */
#define RT_NUM_STAGES 10
#define RT_NUM_CLASSES 5

/* Buffers of the last rt_render_frame, row 0 = bottom. */
#define RT_FRAME_DENOISED 0          /* float4: filtered direct + indirect, alpha 1 */
#define RT_FRAME_DIRECT 1            /* float4: filtered direct, .w variance */
#define RT_FRAME_INDIRECT 2
#define RT_FRAME_HISTORY_DIRECT 3    /* float4: temporal blend (next frame's history), .w variance */
#define RT_FRAME_HISTORY_INDIRECT 4
#define RT_FRAME_MOMENTS_DIRECT 5    /* float2: luminance moments */
#define RT_FRAME_MOMENTS_INDIRECT 6
#define RT_FRAME_HISTORY_LENGTH 7    /* int32 */
#define RT_FRAME_VARIANCE_DIRECT 8   /* float4: variance-pass output */
#define RT_FRAME_VARIANCE_INDIRECT 9
#define RT_FRAME_G_POSITION_TRI 10   /* float4: world position, triangle index within its mesh */
#define RT_FRAME_G_NORMAL_MAT 11     /* uint16x4: normal*0.5+0.5 in 1/65535 units, material id */
#define RT_FRAME_G_BARY_MESH 12      /* uint16x4: barycentrics in 1/65535 units, mesh id + 1 (0: background) */
#define RT_FRAME_G_MOTION_DEPTH 13   /* float4: NDC motion (current - previous), linear depth, fwidth(depth) */

typedef struct rt_scene rt_scene;
typedef struct rt_renderer rt_renderer;

typedef struct rt_load_opts {
    int bvh_builder; /* 0 = automatic (the GPU for scenes of >= 65536 triangles when one is present),
                        1 = host, 2 = GPU; every choice builds the same tree */
} rt_load_opts;

typedef struct rt_scene_info {
    int num_triangles;
    int num_meshes;
    int num_materials;
    int num_lights;
    int num_bvh_nodes;      /* reference-layout node count (interior + leaves) */
    int bvh_max_depth;
    int bvh_fallback_splits; /* splits where the reference build would not terminate */
    int width, height, spp, max_bounces;
    float rr;
    int denoise;
    uint32_t seed;
    int rr_start_depth;
    float bounds_lo[3], bounds_hi[3];
    float camera_position[3], camera_left_down[3], camera_du[3], camera_dv[3];
    int num_textures;       /* entries of the scene's "textures" array */
    float bvh_build_ms;     /* wall time of the BVH build */
    int bvh_on_gpu;         /* 1 when the GPU built it */
    float bvh_gpu_phase_ms[4]; /* GPU build: setup + upload, levels, download, host numbering */
    float load_phase_ms[3];    /* OBJ import (parse, transforms), triangle normals and areas, environment map */
    float refit_ms;            /* wall time of the last rt_scene_refit (0: none yet) */
    int refit_on_gpu;          /* 1 when the last refit ran on the GPU copy of the scene */
    float refit_phase_ms[3];   /* host arrays (copy, normals, areas); uploads; node boxes (kernels or host) */
} rt_scene_info;

typedef struct rt_render_params {
    int max_bounces;        /* < 0: scene value */
    float rr;               /* < 0: scene value (continue probability) */
    int rr_start_depth;     /* < 0: scene value; 0 = roulette from the first bounce (reference) */
    int64_t seed;           /* < 0: scene value */
    int batch_spp;          /* samples per pass; 0 = automatic: the scene's spp (at most 64) when the queue
                               budget holds them (queues are sized for this at creation) */
    int64_t queue_budget_bytes; /* 0 = 16 GiB; a renderer holds 252 B per path of a pass (batch_spp x pixels)
                                   plus 32 B per pixel of accumulators */
    int flush_survivors;    /* 0 = reference: paths alive after max_bounces contribute nothing */
    int profile;            /* 1 = per-stage HIP event timing in rt_stats */
    int env_sampling;       /* 0 = CDF binary search (reference, texture.cu:112-153); 1 = alias tables */
    int disable_graphs;     /* 1 = launch every kernel of a pass; 0 = replay passes as captured HIP graphs */
    int ray_sort;           /* 1 = before each extend, reorder the rays of every queue slice by direction
                               octant and origin cell (38 more bytes per path; same results, faster
                               traversal on coherent scenes such as closed rooms) */
    int shadow_any_hit;     /* 0 = the reference's shadow test (Light::Visible, Src/Core/light.cu:5-27):
                               the closest hit from the light point, then its 0.1 tolerances;
                               1 = the same test with an early exit: a shadow ray stops at the first hit
                               that lies, by its t, more than 0.11 (plus a rounding margin) in front of
                               the shading point on a triangle it does not graze, and counts as occluded
                               -- what the closest-hit test decides whenever the closest hit lies on the
                               ray where its t says; rays without such a hit are traced exactly as in
                               mode 0;
                               2 = fast mode, the segment test: any triangle the reference's ray would
                               meet from the front, between 0.1 past the shading point and 0.01 short
                               of the light point, occludes; the search starts at the shading point
                               and stops at the first such hit.  An approximation: the reference also
                               calls a ray occluded by a hit behind the shading point or at its
                               distance from the light */
} rt_render_params;

typedef struct rt_stats_out {
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t extension_node_visits; /* interior nodes whose child pair was tested (counting renders only) */
    uint64_t extension_tri_tests;
    uint64_t shadow_node_visits;
    uint64_t shadow_tri_tests;
    uint64_t probe_rays;
    double stage_ms[RT_NUM_STAGES];
    int stage_launches[RT_NUM_STAGES];
    int spp_done;
    int batch_spp;
    /* Counting renders only (count_traversal): extension rays that left the scene, shadow
       rays that found their light point visible, and extension rays per shading class of
       what they hit (light, diffuse, GGX, dielectric, subsurface; misses are in no class)
       -- the path counts behind the stages' byte accounting. */
    uint64_t miss_paths;
    uint64_t shadow_lit;
    uint64_t class_paths[RT_NUM_CLASSES];
} rt_stats_out;

/* Orbit camera (CameraParam, Src/Core/camera.h:11-84): the basis follows from these. */
typedef struct rt_camera {
    float target[3];
    float distance;
    float theta_deg; /* polar angle from +y */
    float phi_deg;   /* azimuth about +y, from +z towards +x */
} rt_camera;

/* DenoiseParam (Src/Utils/ui.h:14-25). */
typedef struct rt_denoise_params {
    int max_history;          /* 24 */
    float sigma_light;        /* 8 */
    float sigma_normal;       /* 256 */
    float sigma_depth;        /* 1 */
    float reprojection_depth; /* 0.1 */
    int atrous_iterations;    /* 5 (clamped to 1..8) */
} rt_denoise_params;

/* Fill *p with "use the scene's values" defaults. */
void rt_render_params_default(rt_render_params* p);
/* Fill *p with the reference's DenoiseParam defaults. */
void rt_denoise_params_default(rt_denoise_params* p);

/* Scene::Scene (Src/Core/scene.cpp:50-57). */
int rt_scene_load(const char* json_path, const rt_load_opts* opts, rt_scene** out);
int rt_scene_info_get(const rt_scene* s, rt_scene_info* out);
void rt_scene_destroy(rt_scene* s);
/* Refit: move the scene's vertices without rebuilding its BVH.  No reference
   counterpart: the reference constructs a new Scene (scene.cpp:50-57) and so
   rebuilds the tree.  pos holds 9 floats per triangle in scene order (the order
   rt_scene_get_triangles returns) for all num_triangles triangles of the scene;
   nrm (same layout) replaces the shading normals, NULL keeps them.  Geometric
   normals, areas, area-light tables and every node box follow (a leaf's box is
   the union of its triangles' boxes, an interior node's the union of its
   children's); topology, triangle order, materials and the camera are kept.
   When the scene is resident on the GPU (a renderer, pool or GPU query has
   used it) the refit runs there, in place, after all work on the device has
   finished; renderers built on the scene see the new geometry from their
   next pass, and their accumulated samples are not reset (rt_reset).
   RT_ERR_INVALID_ARG for a wrong count or a non-finite coordinate, before
   anything changes. */
int rt_scene_refit(rt_scene* s, const float* pos, const float* nrm, int64_t num_triangles);

/* Renderer::Renderer (Src/Core/renderer.cuh:1094-1100). */
int rt_renderer_create(rt_scene* s, const rt_render_params* p, rt_renderer** out);
void rt_renderer_destroy(rt_renderer* r);
/* Accumulate spp more samples per pixel (RenderLoop, renderer.cuh:1144-1193). Asynchronous. */
int rt_render(rt_renderer* r, int spp);
/* Same with traversal counters (node visits / triangle tests) enabled. */
int rt_render_counted(rt_renderer* r, int spp);
int rt_synchronize(rt_renderer* r);
/* Clear the accumulators (framebufferReset, renderer.cuh:1121-1124). */
int rt_reset(rt_renderer* r);
int rt_reset_stats(rt_renderer* r);
/* Read back an accumulator: RT_ACCUM_DIRECT, RT_ACCUM_INDIRECT or RT_ACCUM_MEAN. */
int rt_read_accum(rt_renderer* r, int which, float* rgba_out);
/* Gamma-2.2 8-bit colour, top row first (GatherKernel colour mode). */
int rt_read_rgba8(rt_renderer* r, uint8_t* rgba_out);
/* GatherKernel's render-target modes (renderer.cuh:956-1029) as RGBA8, top row first:
   RT_GATHER_COLOR = rt_read_rgba8; DEPTH, NORMAL, ID, POSITION = the debug views of the last
   real-time frame's G-buffer; DENOISED = the colour mode with denoising on (the last frame). */
int rt_read_gather(rt_renderer* r, int mode, uint8_t* rgba_out);
/* Write an image file (UI::SaveScreenshot, Src/Utils/ui.cpp:370-380, writes the gathered colour as
   an RGBA PNG): RT_IMAGE_PNG_COLOR = rt_read_rgba8's pixels, RT_IMAGE_PNG_DENOISED = the same gather
   of the last real-time frame's denoised colour (GatherKernel's denoise branch), RT_IMAGE_PFM_MEAN /
   RT_IMAGE_PFM_DENOISED = the float colour as a PFM. */
int rt_save_image(rt_renderer* r, int which, const char* path);
int rt_stats(rt_renderer* r, rt_stats_out* out);
const char* rt_last_error(void);

/* Current orbit camera of the renderer. */
int rt_get_camera(rt_renderer* r, rt_camera* out);
/* Move the camera; clears the accumulators (a camera move resets the frame buffer). */
int rt_set_camera(rt_renderer* r, const rt_camera* cam);
/* One real-time frame: G-buffer against the previous frame's camera, spp fresh samples
   per pixel (sample indices frame*spp ...), temporal reprojection, variance estimate and
   a-trous filtering (skipped on the first frame after a history reset).  p NULL: defaults.
   Asynchronous; rt_read_accum afterwards returns this frame's raw sums. */
int rt_render_frame(rt_renderer* r, int spp, const rt_denoise_params* p);
/* Forget the temporal history; the next frame is a first frame. */
int rt_reset_history(rt_renderer* r);
/* Copy an RT_FRAME_* buffer of the last frame; bytes must be its exact size
   (width*height times 16 for float4, 8 for float2 / uint16x4, 4 for int32). */
int rt_read_frame(rt_renderer* r, int which, void* out, size_t bytes);
/* The last frame's denoised colour, float RGBA, width*height*4 (rt_read_frame(RT_FRAME_DENOISED)). */
int rt_read_denoised(rt_renderer* r, float* rgba_out);
/* One G-buffer plane of the last frame (channel 0..3 = RT_FRAME_G_POSITION_TRI, _G_NORMAL_MAT,
   _G_BARY_MESH, _G_MOTION_DEPTH), the reference's four render targets (renderer.cuh:1356-1503),
   16 or 8 bytes per pixel as listed with those ids. */
int rt_read_gbuffer(rt_renderer* r, int channel, void* out);

/* ---- pools: many jobs over resident scenes, one wavefront engine ---- */
typedef struct rt_pool rt_pool;

typedef struct rt_pool_params {
    int64_t queue_budget_bytes; /* path queues of the shared engine; 0 = 8 GiB */
    int64_t hbm_budget_bytes;   /* resident scenes + job state admitted at once; 0 = 64 GiB */
    int profile;                /* 1 = per-stage HIP event timing in rt_pool_stats */
    int cohort_ticks;           /* 0 = a tick traces every real-time job's frame, then filters the frames
                                   on three streams in turn; 1 = two cohorts alternate, one traced while
                                   the other's frames are filtered on two side streams (a real-time job
                                   advances every second tick; the faster mode for saturated serving) */
    int max_realtime_jobs;      /* rt_pool_serve's continuous batching admits at most this many real-time
                                   jobs at once (later ones wait in arrival order), which bounds the work of
                                   a tick and so frame latency; 0 = as many as the HBM budget allows */
} rt_pool_params;

typedef struct rt_job_desc {
    int scene;                        /* slot from rt_pool_add_scene */
    int spp;                          /* accumulate: total samples; real-time: samples per frame */
    int frames;                       /* 0 = accumulate job; > 0 = real-time job with SVGF */
    const rt_camera* cameras;         /* real-time: `frames` cameras; accumulate: NULL or one */
    const rt_render_params* render;   /* NULL: the scene's values */
    const rt_denoise_params* denoise; /* NULL: defaults */
} rt_job_desc;

typedef struct rt_job_status {
    int done;
    int spp_done;     /* accumulate: samples so far; real-time: samples of the last frame */
    int frames_done;
    double submit_s, start_s, done_s; /* seconds on the pool's clock; -1 = not yet */
} rt_job_status;

typedef struct rt_serve_request {
    double arrival_s; /* seconds after the start of the run */
    rt_job_desc job;
    float* result_rgba; /* optional (NULL: none), width*height*4 floats, row 0 = bottom: receives the job's
                           result when it completes -- an accumulate job's (direct + indirect) / spp
                           (RT_ACCUM_MEAN), a real-time job's last denoised frame (RT_FRAME_DENOISED) */
} rt_serve_request;

typedef struct rt_serve_report {
    int jobs, frames;
    double makespan_s;   /* first arrival to last completion */
    double jobs_per_s;   /* jobs / makespan */
    double frame_p50_ms, frame_p99_ms, job_p50_ms, job_p99_ms;
} rt_serve_report;

/* Launch shape and pass size of a pool. */
typedef struct rt_pool_tuning {
    float grid_scale;    /* persistent stage grids = occupancy x CUs x scale (1 = default) */
    int64_t pass_paths;  /* paths per pass; 0 = the queue capacity */
    int tile_width;      /* per-pixel frame kernels on tile_width x (256 / tile_width) tiles: 8..128
                            (the first a-trous pass always runs on 16 x 16 LDS tiles) */
    int ray_sort;        /* 1 = reorder extension rays before extend (rt_render_params.ray_sort);
                            38 bytes per path of device memory beyond the queue budget */
} rt_pool_tuning;

typedef struct rt_tune_report {
    rt_pool_tuning best;
    double baseline_ms, best_ms; /* time of the sample jobs with the starting and the chosen tuning */
    int runs;
} rt_tune_report;

int rt_pool_create(const rt_pool_params* p, rt_pool** out);
void rt_pool_destroy(rt_pool* p);
/* The scene must outlive the pool. */
int rt_pool_add_scene(rt_pool* p, rt_scene* s, int* slot_out);
/* Admit a job; RT_ERR_BUSY when its state does not fit the HBM budget now. */
int rt_pool_submit(rt_pool* p, const rt_job_desc* d, int* job_out);
/* One tick over the active jobs (synchronous): each accumulate job that fits gets as many
   samples as fit; real-time jobs are traced or filtered (see serial_ticks).  served_out: jobs
   that progressed (0 = idle). */
int rt_pool_step(rt_pool* p, int* served_out);
int rt_pool_job_status(rt_pool* p, int job, rt_job_status* out);
int rt_pool_read_accum(rt_pool* p, int job, int which, float* rgba_out);
int rt_pool_read_frame(rt_pool* p, int job, int which, void* out, size_t bytes);
int rt_pool_release(rt_pool* p, int job);
int rt_pool_stats(rt_pool* p, rt_stats_out* out);
int rt_pool_reset_stats(rt_pool* p);
int rt_pool_set_tuning(rt_pool* p, const rt_pool_tuning* t);
int rt_pool_get_tuning(rt_pool* p, rt_pool_tuning* out);
/* Autotuner: time the n sample jobs under candidate tunings (grid scale, then pass size, then tile
   shape, then the ray sort; coordinate descent, each candidate run twice) and keep the fastest.
   The pool must be idle. */
int rt_pool_autotune(rt_pool* p, const rt_job_desc* sample, int n, rt_tune_report* out);
/* Replay an arrival stream: batched = 1 serves it with the pool's continuous batching,
   0 with the plain loop (one job at a time, a fresh single-scene renderer each), 2 with the
   plain loop on resident renderers (one single-scene renderer kept per scene, its history
   and accumulation reset for each job).  Frame
   latency = completion - (arrival for a job's first frame, the previous frame's completion
   otherwise); job latency = last completion - arrival.  Optional outputs: frame latencies
   in request order then frame order, job latencies in request order. */
int rt_pool_serve(rt_pool* p, const rt_serve_request* reqs, int n, int batched, rt_serve_report* out,
                  double* frame_latency_ms, double* job_latency_ms);

/* ---- test and inspection hooks (no reference counterpart) ---- */
/* Scene-order triangle soup: pos 9, normals 9, uv 6 floats and mesh id per triangle (any pointer may be NULL). */
int rt_scene_get_triangles(const rt_scene* s, float* pos, float* nrm, float* uv, int32_t* mesh);
/* Per-mesh (material, light, start, count) and area. */
int rt_scene_get_meshes(const rt_scene* s, int32_t* mesh4, float* area);
/* Reference-layout BVH: 8 words per node as in LinearBVHNode (Src/Core/bvh.h:25-33) and the
   ordered-triangle permutation.  from_packed = 1 rebuilds the array from the device layout. */
int rt_scene_get_bvh(const rt_scene* s, int from_packed, void* nodes_out, int32_t* order_out);
/* Decoded diffuse texture `index` (position in the scene's "textures" array) as
   Texture::Load uploads it (Src/Core/texture.cu:207-240): RGBA8, row 0 = top.  *w, *h
   receive its size (0 x 0 for an entry without 'path' or 'name'); rgba may be NULL. */
int rt_scene_get_texture(const rt_scene* s, int index, int* w, int* h, uint8_t* rgba);
/* The scene's materials in "materials" order, as Scene::ParseMaterials fills them
   (Src/Core/scene.cpp:254-395, defaults from Src/Core/material.h:57-71): 24 words per
   material = int32 type (0 light, 1 diffuse, 2 pbr, 3 dielectric, 4 subsurface), basecolor[3],
   emission[3], roughness, metallic, transmission, ior, scatter_distance[3], int32 texture
   (-1: none), int32 valid (0: an entry the reference skips), then this library's random-walk
   subsurface keys: sigma_a[3], sigma_s[3] (per scene unit; defaults 0 and 1), int32 method
   (0 Burley = the reference's, 1 random walk), one padding word.  *n receives the count; out
   may be NULL. */
int rt_scene_get_materials(const rt_scene* s, int* n, void* out);
/* The clearance grid of a random-walk medium (no reference counterpart; the
   random-walk subsurface mode is this library's): *n cells per axis (0 when
   mesh `mesh` is no such medium), cell index floor((p - lo) * inv) in FP32, and
   n^3 floats (x fastest), each a lower bound of the distance from any point of
   the cell to the mesh.  lo, inv and cells may be NULL. */
int rt_scene_get_medium(const rt_scene* s, int mesh, int* n, float* lo, float* inv, float* cells);
/* The scene's decoded environment map as HDRTexture::Load receives it (Src/Core/texture.cu:
   17-25: stbi_loadf's RGB, row 0 = the file's first scanline; a PFM map is read losslessly):
   3 floats per texel.  *w, *h receive its size (0 x 0 without an environment light); rgb may
   be NULL. */
int rt_scene_get_envmap(const rt_scene* s, int* w, int* h, float* rgb);
/* Diffuse-texture albedo on the GPU at n (u, v) pairs (2 floats each), as shading
   reads it (Texture::Sample, Src/Core/texture.cu:242-249): 3 floats per pair. */
int rt_texture_lookup(rt_scene* s, int index, const float* uv, int n, float* rgb_out);
/* Closest hit on the GPU for n rays given as 8 floats (o.xyz, tmin, d.xyz, tmax);
   hits_out: (t, b1, b2, prim-as-int-bits) per ray; counts_out: (node visits, tri tests). */
int rt_trace_rays(rt_scene* s, const float* rays, int n, float* hits_out, uint32_t* counts_out);
/* rt_trace_rays without allocating device memory.  No counterpart in the reference: for callers
   that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  Rays, hits and counts are staged in one fixed device array of the library
   (RT_NOALLOC_MAX_RAYS rays, 56 bytes each, about 14 MiB per GPU, held until the process ends)
   on the current device; more rays per call is RT_ERR_INVALID_ARG.  The scene must already be
   resident on the device (a renderer, pool or rt_trace_rays call has put it there), else
   RT_ERR_INVALID_ARG.  Calls that overlap on one GPU share that array and race. */
#define RT_NOALLOC_MAX_RAYS 262144
int rt_trace_rays_noalloc(rt_scene* s, const float* rays, int n, float* hits_out, uint32_t* counts_out);
/* Device arithmetic check (op: 0 div, 1 sqrt, 2 sin, 3 cos, 4 acos, 5 exp, 6 log, 7 atan2, 8 philox, 9 rcp, 10 pow). */
int rt_device_math(int op, const float* a, const float* b, float* out, int n);

#ifdef __cplusplus
}
#endif

#endif /* RT_AMD_H */
