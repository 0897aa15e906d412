// Host SAH BVH build.  The tree, the triangle order and the node numbering are
// the reference's (so node arrays can be compared element for element); the
// packed 64-byte layout is this library's device format.
#include "bvh.h"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace rt {
namespace {

// 3.40282e+038 as the reference's float bounds initialiser (reference:
// Src/Core/global.h:8 converted to float in Src/Core/bounds.h:12-13).
constexpr float kFloatMax = (float)3.40282e+038;
constexpr int kBuckets = 12;   // reference: Src/Core/bvh.h:40
constexpr int kMaxLeaf = 4;    // reference: Src/Core/scene.cpp:121

inline float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
inline float smax(float a, float b) { return (a < b) ? b : a; }  // std::max

inline Box empty_box() {
    Box b;
    b.lo = v3(kFloatMax, kFloatMax, kFloatMax);
    b.hi = v3(-kFloatMax, -kFloatMax, -kFloatMax);
    return b;
}
// Bounds3D::Expand(Bounds3D) (reference: Src/Core/bounds.h:35-40).
inline Box expand(const Box& a, const Box& o) {
    Box r;
    r.lo = v3(smin(a.lo.x, o.lo.x), smin(a.lo.y, o.lo.y), smin(a.lo.z, o.lo.z));
    r.hi = v3(smax(a.hi.x, o.hi.x), smax(a.hi.y, o.hi.y), smax(a.hi.z, o.hi.z));
    return r;
}
// Bounds3D::Expand(point) via cwiseMin/cwiseMax (reference: Src/Core/bounds.h:28-33).
inline Box expand_pt(const Box& a, vec3 p) {
    Box r;
    r.lo = v3(smin(a.lo.x, p.x), smin(a.lo.y, p.y), smin(a.lo.z, p.z));
    r.hi = v3(smax(a.hi.x, p.x), smax(a.hi.y, p.y), smax(a.hi.z, p.z));
    return r;
}
// Bounds3D::SurfaceArea (reference: Src/Core/bounds.h:81-87).
inline float surface_area(const Box& b) {
    if (b.lo.x > b.hi.x || b.lo.y > b.hi.y || b.lo.z > b.hi.z) return 0.f;
    vec3 e = b.hi - b.lo;
    return 2.f * ((e.x * e.y + e.y * e.z) + e.z * e.x);
}
// First index of the largest extent (Eigen maxCoeff(&i), strict '>').
inline int max_extent(const Box& b) {
    vec3 e = b.hi - b.lo;
    int axis = 0;
    float m = e.x;
    if (e.y > m) { m = e.y; axis = 1; }
    if (e.z > m) { axis = 2; }
    return axis;
}
inline float comp(vec3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
// Bounds3D::Offset along one axis (reference: Src/Core/bounds.h:89-97).
inline float offset_axis(const Box& cb, vec3 c, int axis) {
    float o = comp(c, axis) - comp(cb.lo, axis);
    float ext = comp(cb.hi, axis) - comp(cb.lo, axis);
    if (ext > 0) o /= ext;
    return o;
}
inline int bucket_of(const Box& cb, vec3 c, int axis) {
    int b = (int)((float)kBuckets * offset_axis(cb, c, axis));
    if (b == kBuckets) b = kBuckets - 1;
    return b;
}

struct Builder {
    const std::vector<Box>& tb;      // per-triangle bounds
    const std::vector<vec3>& cen;    // per-triangle centroid
    std::vector<int32_t>& idx;       // working permutation (partitioned in place)
    BvhBuild& out;

    int build(int start, int end, int depth) {
        int self = (int)out.nodes.size();
        out.nodes.push_back(LinearNode{-1, -1, {0, 0, 0}, {0, 0, 0}});
        if (depth > out.max_depth) out.max_depth = depth;

        Box bounds = empty_box();
        for (int i = start; i < end; ++i) bounds = expand(bounds, tb[idx[i]]);
        set_box(self, bounds);

        if (end - start <= kMaxLeaf) {
            make_leaf(self, start, end);
            return self;
        }

        Box cb = empty_box();
        for (int i = start; i < end; ++i) cb = expand_pt(cb, cen[idx[i]]);
        int axis = max_extent(cb);

        int counts[kBuckets] = {0};
        Box bb[kBuckets];
        for (int b = 0; b < kBuckets; ++b) bb[b] = empty_box();
        for (int i = start; i < end; ++i) {
            int b = bucket_of(cb, cen[idx[i]], axis);
            counts[b]++;
            bb[b] = expand(bb[b], tb[idx[i]]);
        }

        // Costs for splits after bucket i, i in [0, 10]; the right side sums
        // buckets i+1..10 (bucket 11 is not included), and the search starts
        // at i = 1 (reference: Src/Core/bvh.cu:117-144).
        const int nsplit = kBuckets - 1;
        float cost[kBuckets - 1];
        float sa = surface_area(bounds);
        for (int i = 0; i < nsplit; ++i) {
            Box b0 = empty_box(), b1 = empty_box();
            int c0 = 0, c1 = 0;
            for (int j = 0; j <= i; ++j) { b0 = expand(bb[j], b0); c0 += counts[j]; }
            for (int j = i + 1; j < nsplit; ++j) { b1 = expand(bb[j], b1); c1 += counts[j]; }
            cost[i] = .125f + ((float)c0 * surface_area(b0) + (float)c1 * surface_area(b1)) / sa;
        }
        float min_cost = kFloatMax;
        int best = -1;
        for (int i = 1; i < nsplit; ++i) {
            if (cost[i] < min_cost) { min_cost = cost[i]; best = i; }
        }

        int mid = -1;
        if (best >= 0) mid = partition(start, end, cb, axis, best);
        if (mid <= start || mid >= end) {
            // The reference recurses on the same range forever here (all
            // centroids in one bucket); split the range in half instead.
            out.fallback_splits++;
            mid = (start + end) / 2;
        }
        build(start, mid, depth + 1);
        int right = build(mid, end, depth + 1);
        out.nodes[self].tri_index = -1;
        out.nodes[self].second = right;
        return self;
    }

    // libstdc++ std::partition on a bidirectional range (two-pointer swap),
    // predicate "bucket <= best" (reference: Src/Core/bvh.cu:159-165).
    int partition(int first, int last, const Box& cb, int axis, int best) {
        auto pred = [&](int p) { return bucket_of(cb, cen[idx[p]], axis) <= best; };
        while (true) {
            while (true) {
                if (first == last) return first;
                if (pred(first)) ++first;
                else break;
            }
            --last;
            while (true) {
                if (first == last) return first;
                if (!pred(last)) --last;
                else break;
            }
            std::swap(idx[first], idx[last]);
            ++first;
        }
    }

    void set_box(int n, const Box& b) {
        LinearNode& ln = out.nodes[n];
        ln.lo[0] = b.lo.x; ln.lo[1] = b.lo.y; ln.lo[2] = b.lo.z;
        ln.hi[0] = b.hi.x; ln.hi[1] = b.hi.y; ln.hi[2] = b.hi.z;
    }
    void make_leaf(int n, int start, int end) {
        out.nodes[n].tri_index = (int32_t)out.order.size();
        out.nodes[n].second = end - start;
        for (int i = start; i < end; ++i) out.order.push_back(idx[i]);
    }
};

Box node_box(const LinearNode& n) {
    Box b;
    b.lo = v3(n.lo[0], n.lo[1], n.lo[2]);
    b.hi = v3(n.hi[0], n.hi[1], n.hi[2]);
    return b;
}

}  // namespace

void build_bvh(const float* pos, int num_tris, BvhBuild& out) {
    out = BvhBuild();
    if (num_tris <= 0) return;
    std::vector<Box> tb(num_tris);
    std::vector<vec3> cen(num_tris);
    for (int t = 0; t < num_tris; ++t) {
        const float* p = pos + 9 * (size_t)t;
        vec3 a = v3(p[0], p[1], p[2]), b = v3(p[3], p[4], p[5]), c = v3(p[6], p[7], p[8]);
        // Bounds3D(v0, v1, v2): Min(v0, Min(v1, v2)) (reference: Src/Core/bounds.h:22-26).
        Box bx;
        bx.lo = v3(smin(a.x, smin(b.x, c.x)), smin(a.y, smin(b.y, c.y)), smin(a.z, smin(b.z, c.z)));
        bx.hi = v3(smax(a.x, smax(b.x, c.x)), smax(a.y, smax(b.y, c.y)), smax(a.z, smax(b.z, c.z)));
        tb[t] = bx;
        cen[t] = (bx.lo + bx.hi) * 0.5f;
    }
    std::vector<int32_t> idx(num_tris);
    for (int t = 0; t < num_tris; ++t) idx[t] = t;
    out.order.reserve(num_tris);
    out.nodes.reserve(2 * (size_t)num_tris);
    Builder bld{tb, cen, idx, out};
    bld.build(0, num_tris, 0);
    out.bounds = node_box(out.nodes[0]);

    // Pack interior nodes (DFS order) into 64-byte child-pair lines.
    const int nn = (int)out.nodes.size();
    std::vector<int32_t> compact(nn, -1);
    int ni = 0;
    for (int i = 0; i < nn; ++i)
        if (out.nodes[i].tri_index == -1) compact[i] = ni++;
    auto ref_of = [&](int i) -> uint32_t {
        const LinearNode& n = out.nodes[i];
        if (n.tri_index == -1) return (uint32_t)compact[i];
        return leaf_ref((uint32_t)n.tri_index, (uint32_t)n.second);
    };
    out.packed.resize(ni);
    for (int i = 0; i < nn; ++i) {
        if (out.nodes[i].tri_index != -1) continue;
        const LinearNode& L = out.nodes[i + 1];
        const LinearNode& R = out.nodes[out.nodes[i].second];
        PackedNode& p = out.packed[compact[i]];
        p.w0 = make_float4(L.lo[0], L.lo[1], L.lo[2], L.hi[0]);
        p.w1 = make_float4(L.hi[1], L.hi[2], R.lo[0], R.lo[1]);
        p.w2 = make_float4(R.lo[2], R.hi[0], R.hi[1], R.hi[2]);
        p.w3 = make_int4((int)ref_of(i + 1), (int)ref_of(out.nodes[i].second), i, 0);
    }
    out.root_ref = ref_of(0);
    out.max_stack = out.max_depth + 1;
}

void refit_bvh(const float* pos, int num_tris, BvhBuild& b) {
    const int nn = (int)b.nodes.size();
    if (nn == 0) return;
    if ((int)b.order.size() != num_tris) throw std::runtime_error("refit: triangle count differs from the tree's");
    // Children follow their parent in DFS pre-order, so one reverse sweep
    // sees every child before its parent.
    for (int i = nn - 1; i >= 0; --i) {
        LinearNode& n = b.nodes[i];
        Box box = empty_box();
        if (n.tri_index >= 0) {
            for (int k = 0; k < n.second; ++k) {
                const float* p = pos + 9 * (size_t)b.order[(size_t)n.tri_index + k];
                Box t;
                t.lo = v3(smin(p[0], smin(p[3], p[6])), smin(p[1], smin(p[4], p[7])), smin(p[2], smin(p[5], p[8])));
                t.hi = v3(smax(p[0], smax(p[3], p[6])), smax(p[1], smax(p[4], p[7])), smax(p[2], smax(p[5], p[8])));
                box = expand(box, t);
            }
        } else {
            box = expand(expand(box, node_box(b.nodes[i + 1])), node_box(b.nodes[n.second]));
        }
        n.lo[0] = box.lo.x; n.lo[1] = box.lo.y; n.lo[2] = box.lo.z;
        n.hi[0] = box.hi.x; n.hi[1] = box.hi.y; n.hi[2] = box.hi.z;
    }
    for (PackedNode& p : b.packed) {
        const int i = p.w3.z;
        const LinearNode& L = b.nodes[i + 1];
        const LinearNode& R = b.nodes[b.nodes[i].second];
        p.w0 = make_float4(L.lo[0], L.lo[1], L.lo[2], L.hi[0]);
        p.w1 = make_float4(L.hi[1], L.hi[2], R.lo[0], R.lo[1]);
        p.w2 = make_float4(R.lo[2], R.hi[0], R.hi[1], R.hi[2]);
    }
    b.bounds = node_box(b.nodes[0]);
}

namespace {
void unpack_rec(const std::vector<PackedNode>& packed, uint32_t ref, const float lo[3], const float hi[3],
                std::vector<LinearNode>& out) {
    int self = (int)out.size();
    LinearNode ln;
    for (int k = 0; k < 3; ++k) { ln.lo[k] = lo[k]; ln.hi[k] = hi[k]; }
    if (ref & kLeafBit) {
        ln.tri_index = (int32_t)((ref & ~kLeafBit) >> 3);
        ln.second = (int32_t)(ref & 7u);
        out.push_back(ln);
        return;
    }
    ln.tri_index = -1;
    ln.second = -1;
    out.push_back(ln);
    const PackedNode& p = packed[ref];
    float llo[3] = {p.w0.x, p.w0.y, p.w0.z}, lhi[3] = {p.w0.w, p.w1.x, p.w1.y};
    float rlo[3] = {p.w1.z, p.w1.w, p.w2.x}, rhi[3] = {p.w2.y, p.w2.z, p.w2.w};
    unpack_rec(packed, (uint32_t)p.w3.x, llo, lhi, out);
    out[self].second = (int32_t)out.size();
    unpack_rec(packed, (uint32_t)p.w3.y, rlo, rhi, out);
}
}  // namespace

void unpack_bvh(const std::vector<PackedNode>& packed, uint32_t root_ref, const Box& root_box,
                std::vector<LinearNode>& out) {
    out.clear();
    float lo[3] = {root_box.lo.x, root_box.lo.y, root_box.lo.z};
    float hi[3] = {root_box.hi.x, root_box.hi.y, root_box.hi.z};
    unpack_rec(packed, root_ref, lo, hi, out);
}

}  // namespace rt
