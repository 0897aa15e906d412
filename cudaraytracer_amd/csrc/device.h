// Device-side data layout of a loaded scene and of the wavefront queues.
#pragma once

#include "bvh.h"
#include "rt_math.h"
#include "scene.h"

namespace rt {

// Scene arrays in HBM.  Triangles are stored twice, as the reference does
// (once in BVH order for traversal and shading, once in scene order for the
// light-triangle table used by area-light sampling).
struct DevScene {
    const PackedNode* nodes;      // interior nodes, 64 B each
    uint32_t root_ref;
    const float4* tri_isect;      // 3 x float4 per triangle (BVH order): v0, e1 = v1-v0, e2 = v2-v0
    const float4* tri_attr;       // 7 x float4 per triangle (BVH order), see kAttr* below
    const int2* mesh_info;        // (material, light) per mesh
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
};

// tri_attr layout, 7 float4 per triangle:
//   [0] v0.xyz, mesh id (int bits)     [1] v1.xyz, gn.x
//   [2] v2.xyz, gn.y                   [3] n0.xyz, gn.z
//   [4] n1.xyz, uv0.x                  [5] n2.xyz, uv0.y
//   [6] uv1.xy, uv2.xy
constexpr int kAttrStride = 7;

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

// Counter slots (int32) in device memory.  No memset is used: kernels zero
// the slots the next kernel needs, so the bounce loop can be graph-captured.
enum CounterSlot : int {
    kCntQueue0 = 0,
    kCntQueue1 = 1,
    kCntShadow = 2,
    kFetchExtend = 3,
    kFetchShade = 4,
    kFetchShadow = 5,
    kNumCounters = 8,
};
// Statistics slots (uint64).
enum StatSlot : int {
    kStatExtRays = 0,
    kStatShadowRays = 1,
    kStatExtNodes = 2,
    kStatExtTris = 3,
    kStatShadowNodes = 4,
    kStatShadowTris = 5,
    kStatProbeRays = 6,
    kNumStats = 8,
};

}  // namespace rt
