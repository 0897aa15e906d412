// Host SAH BVH builder and the packed device node layout.
#pragma once

#include <stdint.h>
#include <vector>

#include "rt_math.h"

namespace rt {

// Axis-aligned box.  The empty box is (+FLOATMAX, -FLOATMAX) with FLOATMAX the
// reference's 3.40282e+38 rounded to float (reference: Src/Core/global.h:8,
// Src/Core/bounds.h:10-14).
struct Box {
    vec3 lo, hi;
};

// One node of the reference's flattened DFS pre-order array
// (reference: Src/Core/bvh.h:25-33): tri_index == -1 marks an interior node
// whose left child is the next node and whose right child is `second`; a leaf
// has tri_index = first ordered triangle and second = triangle count.
struct LinearNode {
    int32_t tri_index;
    int32_t second;
    float lo[3];
    float hi[3];
};
static_assert(sizeof(LinearNode) == 32, "reference node is 32 bytes");

// Device node: one 64-byte line per interior node holding both children's
// boxes and child references, so one traversal step is one line fetch.
//   w0 = (L.lo.x, L.lo.y, L.lo.z, L.hi.x)
//   w1 = (L.hi.y, L.hi.z, R.lo.x, R.lo.y)
//   w2 = (R.lo.z, R.hi.x, R.hi.y, R.hi.z)
//   w3 = (ref L, ref R, reference node id, 0) as int bits
// A child reference is either the index of an interior node (bit 31 clear) or
// a leaf: bit 31 set, bits 3..30 first ordered triangle, bits 0..2 count.
struct PackedNode {
    float4 w0, w1, w2;
    int4 w3;
};
static_assert(sizeof(PackedNode) == 64, "packed node is one 64-byte line");

constexpr uint32_t kLeafBit = 0x80000000u;
RT_HD uint32_t leaf_ref(uint32_t first, uint32_t count) { return kLeafBit | (first << 3) | count; }

struct BvhBuild {
    std::vector<LinearNode> nodes;       // reference layout, DFS pre-order
    std::vector<int32_t> order;          // ordered position -> original triangle index
    std::vector<PackedNode> packed;      // interior nodes only, DFS order
    uint32_t root_ref = 0;               // leaf ref when the whole scene is one leaf
    int max_depth = 0;
    int max_stack = 0;                   // deepest traversal stack the tree can need
    int fallback_splits = 0;             // splits where the reference would not terminate
    Box bounds;
    // GPU build phases (ms): setup + upload, levels, download, host numbering.
    double gpu_phase_ms[4] = {0, 0, 0, 0};
};

// Binned SAH build (12 buckets, leaves of <= 4 triangles) restating
// reference Src/Core/bvh.cu:44-175 and the flatten at bvh.cu:186-206.
// `pos` holds 9 floats (v0, v1, v2) per triangle in scene order.
void build_bvh(const float* pos, int num_tris, BvhBuild& out);
// The same build on the current GPU (bvh_gpu.hip): an identical BvhBuild.
void build_bvh_gpu(const float* pos, int num_tris, BvhBuild& out);

// Refit: new vertex positions (9 floats per triangle, scene order) for the
// same triangles over the same tree.  A leaf's box is the union of its
// triangles' boxes in leaf order, starting from the empty box (as the build
// forms a node's box); an interior node's box is its left child's box
// expanded by its right child's.  Node boxes, packed boxes and `bounds` are
// rewritten; the topology, the triangle order and the numbering are kept.
// The reference has no refit: it rebuilds the whole scene.
void refit_bvh(const float* pos, int num_tris, BvhBuild& b);

// Rebuild the reference LinearNode array from the packed form (test hook).
void unpack_bvh(const std::vector<PackedNode>& packed, uint32_t root_ref, const Box& root_box,
                std::vector<LinearNode>& out);

}  // namespace rt
