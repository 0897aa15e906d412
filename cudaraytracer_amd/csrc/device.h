// Device-side data layout of a loaded scene and of the wavefront queues.
#pragma once

#include "bvh.h"
#include "rt_math.h"
#include "scene.h"

namespace rt {

// Scene arrays in HBM.  Triangles are stored twice, as the reference does
// (once in BVH order for traversal and shading, once in scene order for the
// light-triangle table used by area-light sampling).
// Global-memory views of the scene arrays.  Their pointers reach the kernels
// inside structs, where the compiler cannot tell global memory from LDS and
// emits flat loads with combined waits; these casts let it issue global
// loads.  float4 goes through the native vector type, which (unlike the
// float4 class) copies out of a qualified address space in both the host and
// the device compilation pass.  Device code only.
typedef float NativeF4 __attribute__((ext_vector_type(4)));
using GlobalV4 = const __attribute__((address_space(1))) NativeF4;
template <typename T>
using Global = const __attribute__((address_space(1))) T;
RT_HD float4 as_f4(NativeF4 v) { return make_float4(v.x, v.y, v.z, v.w); }
RT_HD float4 ldg4(const float4* p) { return as_f4(*(GlobalV4*)p); }
template <typename T>
RT_HD Global<T>& gref(const T* p) { return *(Global<T>*)p; }

struct DevScene {
    const PackedNode* nodes;      // interior nodes, 64 B each
    uint32_t root_ref;
    const float4* tri_isect;      // 3 x float4 per triangle (BVH order): v0, e1 = v1-v0, e2 = v2-v0
    const float4* tri_attr;       // 8 x float4 per triangle (BVH order), see below
    const MaterialRec* materials;
    const LightRec* lights;
    int num_lights;
    int env_light;
    const float4* light_tris;     // 4 x float4 per light triangle: v0, v1, v2, geometric normal
    vec3 bounds_lo, bounds_hi;
    // Environment map (RGB float, row 0 = top) and its sampling tables.
    const float* env_rgb;
    int env_w, env_h;
    const float* env_row_cdf;     // env_h * env_w, per-row conditional CDF
    const float* env_marg_cdf;    // env_h, marginal CDF over rows
    const float* env_pdf;         // env_h * env_w, pdf in uv-space
    // Alias tables over the same texel distribution (Walker/Vose): per row a
    // conditional table of env_w entries, and a marginal table of env_h
    // entries; entry = (acceptance threshold, alias index bits).
    const float2* env_alias_row;
    const float2* env_alias_marg;
    int env_alias;                // per job: 1 = sample the environment with the alias tables
    // Diffuse textures: RGBA8 texels of all textures back to back (row 0 =
    // top), and per texture (first texel, width, height, 0).
    const uint32_t* tex_texels;
    const int32_t* tex_info;
    // Random-walk media: per mesh the root reference of its own tree
    // (kNoMedium for other meshes), the trees' nodes and triangle records
    // (all media back to back, references absolute), and per medium triangle
    // its position in the scene's BVH order (for its attribute line).
    const uint32_t* media_root;
    const PackedNode* media_nodes;
    const float4* media_isect;
    const int32_t* media_prim;
    // Per mesh two float4 of its medium's clearance grid: (lo.xyz, n as int
    // bits), (inv cell size.xyz, first cell as int bits); and the cells.
    const float4* media_grid;
    const float* media_clear;
};
constexpr uint32_t kNoMedium = 0xffffffffu;

// tri_attr layout, 8 float4 = one 128-byte line per triangle:
//   [0] v0.xyz, index within its mesh  [1] v1.xyz, gn.x
//   [2] v2.xyz, gn.y                   [3] n0.xyz, gn.z
//   [4] n1.xyz, uv0.x                  [5] n2.xyz, uv0.y
//   [6] uv1.xy, uv2.xy                 [7] mesh, material, light, shading class (int bits)
// Material and light ids ride in the triangle's line so shading needs no
// mesh-table lookup between the hit and the material.
constexpr int kAttrStride = 8;

// Shading classes: extend bins every path by the class of what it hit, and
// one specialised shade kernel per class runs without cross-material divergence.
enum ShadeClass : int {
    kClsTerminal = 0,   // light hit: the path ends (a miss ends in the extend stage)
    kClsDiffuse = 1,
    kClsPbr = 2,
    kClsDielectric = 3,
    kClsSubsurface = 4,
    kNumClasses = 5,
};

// Path flags kept in the .w lane of the direction word.
constexpr uint32_t kFlagFirst = 1u;
constexpr uint32_t kFlagSpecular = 2u;

// Per-frame constants of a render pass.
struct FrameParams {
    vec3 cam_pos, left_down, du, dv;
    int width, height, wh;
    uint32_t seed;
    float rr;
    int rr_start;
    int max_bounces;
    uint32_t sample0;     // first sample index of this batch
    int n_paths;          // width*height*samples in this batch
    int bounce;           // current bounce (set per launch)
};

// A render pass carries paths of one or more jobs.  The job's index in the
// pass table rides in the path flags word above the flag bits and in the
// shadow record, so every stage finds its scene and frame constants without a
// search; one job's paths are contiguous in the pass (path index p,
// local index p - path_base).
constexpr int kJobShift = 8;
constexpr int kMaxPassJobs = 256;

struct JobDev {
    DevScene sc;
    FrameParams fp;        // fp.n_paths = paths of this job in the pass, fp.sample0 its first sample
    int path_base;         // first pass path index of the job
    int pix_base;          // first pass pixel index of the job (resolve)
    int flush_survivors;
    int pad;
    float4* acc_d;         // the job's accumulators, W*H each
    float4* acc_i;
};

struct PassHeader {
    int n_jobs;
    int n_paths;
    int n_pixels;
    int region;  // queue entries per slice in this pass: ceil(n_paths / kSlices)
};
// Device pass table: header, then the jobs, 256-byte aligned.
constexpr size_t kPassHeaderBytes = 256;
constexpr size_t kPassTableBytes = kPassHeaderBytes + sizeof(JobDev) * kMaxPassJobs;

// Work-queue counters.  Every queue is split into kSlices slices; each slice
// has its own counter of every kind on its own 256-byte line, because one hot
// atomic address serialises the whole chip (about 11 ns per atomic on MI355X).
// Counter (kind, slice) lives at int offset (kind * kSlices + slice) * kCounterStride.
// No memset is used: kernels zero the counters the next kernel needs, so the
// bounce loop can be graph-captured.
constexpr int kSlices = 64;
constexpr int kCounterStride = 64;
enum CounterKind : int {
    kCntQueue0 = 0,     // live paths in queue 0 / 1 (ping-pong)
    kCntQueue1 = 1,
    kCntShadow = 2,     // shadow records
    kFetchExtend = 3,   // fetch cursors
    kFetchShadow = 4,
    kCntClass0 = 5,     // kNumClasses per-class list lengths
    kFetchClass0 = 10,  // kNumClasses per-class fetch cursors
    kCntCont = 15,      // suspended extension rays stored for the continuation launch
    kFetchCont = 16,    // and its fetch cursor
    kCntShadowCont = 17,  // suspended shadow rays, and their fetch cursor
    kFetchShadowCont = 18,
    kNumCounterKinds = 19,
};
constexpr size_t kCounterInts = (size_t)kNumCounterKinds * kSlices * kCounterStride;

// Statistics slots (uint64).
enum StatSlot : int {
    kStatExtRays = 0,
    kStatShadowRays = 1,
    kStatExtNodes = 2,
    kStatExtTris = 3,
    kStatShadowNodes = 4,
    kStatShadowTris = 5,
    kStatProbeRays = 6,
    kStatMisses = 7,       // extension rays that left the scene (counting renders only)
    kStatShadowLit = 8,    // shadow rays that found their light point visible (counting renders only)
    kStatClassPaths = 9,   // + class: extension rays binned to each shading class (counting renders only)
    kNumStats = kStatClassPaths + kNumClasses,
};

}  // namespace rt
