// Closest-hit BVH traversal for gfx950: 64-byte child-pair nodes, a per-lane
// stack column in LDS, nearest child first.  The visiting order, the slab test
// and the triangle test are the reference's (Src/Core/bvh.cu:253-306,
// Src/Core/bounds.h:60-78, Src/Core/triangle.cu:102-133) so hits are
// bit-identical to a host run of the same FP32 code.
#pragma once

#include "bvh.h"
#include "device.h"
#include "rt_math.h"

namespace rt {

struct RayInv {
    vec3 o, d, inv;
    bool nx, ny, nz;
};

RT_HD RayInv make_ray(vec3 o, vec3 d) {
    RayInv r;
    r.o = o;
    r.d = d;
    r.inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    r.nx = r.inv.x < 0;
    r.ny = r.inv.y < 0;
    r.nz = r.inv.z < 0;
    return r;
}

// Slab test with the reference's 1e-2 slack: (double)x + 1e-2 >= 0 holds
// exactly when x >= -0.01f for every float x, so the test stays in FP32.
// "exit >= -0.01 and exit >= tmin" is one comparison against the larger of
// the two (a loop invariant); a NaN tmin keeps both tests false.
// OCT >= 0 fixes the signs of the inverse direction at compile time (bit k
// set: axis k negative) for rays known to have them; the arithmetic is the
// same, only the per-axis selects disappear.
template <int OCT = -1>
RT_HD bool box_hit(const RayInv& r, float lx, float ly, float lz, float hx, float hy, float hz, float tmin,
                   float tmax, float& t) {
    const bool nx = OCT < 0 ? r.nx : (OCT & 1) != 0;
    const bool ny = OCT < 0 ? r.ny : (OCT & 2) != 0;
    const bool nz = OCT < 0 ? r.nz : (OCT & 4) != 0;
    float ex = (lx - r.o.x) * r.inv.x, xx = (hx - r.o.x) * r.inv.x;
    float ey = (ly - r.o.y) * r.inv.y, xy = (hy - r.o.y) * r.inv.y;
    float ez = (lz - r.o.z) * r.inv.z, xz = (hz - r.o.z) * r.inv.z;
    float ent_x = nx ? xx : ex, ext_x = nx ? ex : xx;
    float ent_y = ny ? xy : ey, ext_y = ny ? ey : xy;
    float ent_z = nz ? xz : ez, ext_z = nz ? ez : xz;
    float exit = fminf(ext_x, fminf(ext_y, ext_z));
    float enter = fmaxf(ent_x, fmaxf(ent_y, ent_z));
    t = enter;
    const float exit_lo = (tmin >= -0.01f || tmin != tmin) ? tmin : -0.01f;
    return (exit - enter >= -0.01f) && (exit >= exit_lo) && (enter <= tmax);
}

// Moller-Trumbore with the triangle's own [0, 1e8] range.  Returns true and
// updates (t, b1, b2) on a closer hit.  CULL rejects back faces as the path
// tracer does (reference: Src/Core/triangle.cu:102-133); without it both
// faces count, as in the rasterised G-buffer (no GL_CULL_FACE).
template <bool CULL = true>
RT_HD bool tri_hit(const RayInv& r, float4 a, float4 e1f, float4 e2f, float& hit_t, float& hb1, float& hb2) {
    vec3 v0 = v3(a.x, a.y, a.z), e1 = v3(e1f.x, e1f.y, e1f.z), e2 = v3(e2f.x, e2f.y, e2f.z);
    vec3 s = r.o - v0;
    vec3 s1 = cross(r.d, e2);
    vec3 s2 = cross(s, e1);
    float det = dot(e1, s1);
    if (CULL ? det <= 0.f : det == 0.f) return false;
    float inv = 1.f / det;
    float t = inv * dot(s2, e2);
    if (t >= hit_t) return false;
    float b1 = inv * dot(s1, s);
    if (b1 < 0.f || b1 > 1.f) return false;
    float b2 = inv * dot(s2, r.d);
    if (b2 < 0.f || b1 + b2 > 1.f) return false;
    if (t < 0.f || t > 1e8f) return false;
    hit_t = t;
    hb1 = b1;
    hb2 = b2;
    return true;
}

struct Hit {
    float t, b1, b2;
    int prim;
};

// Traversal counters for the algorithmic-bytes model (interior nodes whose
// child pair is tested, triangle tests).
struct TravCount {
    uint32_t nodes, tris;
};

// The stack column is LDS; node and triangle loads use the global views of
// device.h, so the loop issues global and LDS instructions, not flat ones.
using LdsU32 = __attribute__((address_space(3))) uint32_t;

// `stk` points at this lane's stack column in LDS; entries are `stride` words apart.
// OCT >= 0: the ray's inverse-direction signs (see box_hit), known for the whole wave.
// EARLY: stop at the first accepted triangle whose t is below t_exit and that
// the ray does not graze (|cos| between ray and triangle normal >= 1e-3); the
// hit then has prim = kEarlyExit.  Without such a triangle the result is the
// closest hit, as with EARLY off.
//
// Each loop iteration is one interior node or ONE triangle of the current
// leaf: a lane in a leaf keeps its next and end triangle in registers and
// tests them over successive iterations.  Every ray makes the reference's
// sequence of tests (a leaf's triangles in order, then the pop), so the hit
// is the same; but a wave no longer runs a whole leaf's triangle loop for the
// few lanes that reached leaves while the others wait, which made the
// triangle tests most of the instructions issued.
constexpr int kEarlyExit = -2;

// Where a closest-hit search stands between two loop iterations: the hit so
// far, the next item to visit (an interior node, or a leaf reference whose
// first triangle and count name the leaf's triangles still to be tested) and
// the number of entries on the lane's stack.  A search started from this
// state makes the same tests in the same order as one that never stopped, so
// a search can be suspended, its state and stack stored, and resumed later on
// another lane of another launch with the same result.
struct TravState {
    Hit h;
    uint32_t ref;
    int sp;
};
// A suspended search keeps at most this many stack entries (deeper lanes do
// not suspend).
constexpr int kSuspendMaxStack = 16;

// The tree is given by its packed nodes, its triangle records and the state
// to start from (see trace_closest_tree below for a search from the root).
// SUSPEND > 0: at the end of an iteration, when fewer than SUSPEND lanes of
// the wave are still searching, every searching lane whose stack holds at
// most kSuspendMaxStack entries stops, leaves its state in `st` (its stack
// stays in the column) and the function returns true; the wave's other
// lanes finished long ago, so the stragglers can be gathered into full waves.
template <bool COUNT, bool CULL = true, int OCT = -1, bool EARLY = false, int SUSPEND = 0>
__device__ __forceinline__ bool trace_closest_run(const PackedNode* tree_nodes, const float4* tree_tris,
                                                  const RayInv& r, float tmin, uint32_t* stk_generic, int stride,
                                                  TravCount& cnt, TravState& st, float t_exit = 0.f) {
    LdsU32* stk = (LdsU32*)stk_generic;
    GlobalV4* nodes = (GlobalV4*)tree_nodes;  // 4 vectors per node
    GlobalV4* tris = (GlobalV4*)tree_tris;
    Hit h = st.h;
    uint32_t ref = st.ref;
    // In a leaf: triangles tn .. te-1 are still to be tested.
    bool leaf = (ref & kLeafBit) != 0;
    uint32_t tn = (ref & ~kLeafBit) >> 3, te = tn + (ref & 7u);
    int sp = st.sp;
    while (true) {
        uint32_t next;
        if (leaf) {
            if (tn < te) {
                GlobalV4* tp = tris + 3 * (size_t)tn;
                float4 a = as_f4(tp[0]), e1 = as_f4(tp[1]), e2 = as_f4(tp[2]);
                if (COUNT) cnt.tris++;
                if (tri_hit<CULL>(r, a, e1, e2, h.t, h.b1, h.b2)) {
                    h.prim = (int)tn;
                    if (EARLY && h.t < t_exit) {
                        vec3 ve1 = v3(e1.x, e1.y, e1.z), ve2 = v3(e2.x, e2.y, e2.z);
                        float det = dot(ve1, cross(r.d, ve2));
                        vec3 n = cross(ve1, ve2);
                        if (det * det >= 1e-6f * dot(n, n)) {
                            h.prim = kEarlyExit;
                            break;
                        }
                    }
                }
                ++tn;
            }
            if (tn < te) goto latch;
            if (sp == 0) break;
            next = stk[--sp * stride];
        } else {
            GlobalV4* np = nodes + 4 * (size_t)ref;
            float4 w0 = as_f4(np[0]), w1 = as_f4(np[1]), w2 = as_f4(np[2]);
            const NativeF4 w3f = np[3];
            const uint32_t rl = (uint32_t)f_as_i(w3f.x), rr = (uint32_t)f_as_i(w3f.y);
            if (COUNT) cnt.nodes++;
            float tr, tl;
            bool hr = box_hit<OCT>(r, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w, tmin, h.t, tr);
            bool hl = box_hit<OCT>(r, w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, tmin, h.t, tl);
            if (hr && hl) {
                // Nearer child next; ties visit the left child first.
                stk[sp++ * stride] = tr < tl ? rl : rr;
                next = tr < tl ? rr : rl;
            } else if (hr || hl) {
                next = hr ? rr : rl;
            } else {
                if (sp == 0) break;
                next = stk[--sp * stride];
            }
        }
        leaf = (next & kLeafBit) != 0;
        ref = next;
        tn = (next & ~kLeafBit) >> 3;
        te = tn + (next & 7u);
    latch:
        // The check sits on the loop's back edge, where the state is the next
        // iteration's, so stopping costs no copies of it.
        // A wave-uniform test: all searching lanes leave the loop together.
        if (SUSPEND > 0 && __builtin_amdgcn_readfirstlane(__builtin_popcountll(__ballot(1))) < SUSPEND) {
            st.h = h;
            st.ref = leaf ? leaf_ref(tn, te - tn) : ref;
            st.sp = sp;
            if (sp <= kSuspendMaxStack) return true;
            // Too deep a stack to store: finish here.
            return trace_closest_run<COUNT, CULL, OCT, EARLY, 0>(tree_nodes, tree_tris, r, tmin, stk_generic, stride,
                                                                 cnt, st, t_exit);
        }
    }
    st.h = h;
    return false;
}

// A search from the root of a tree (the scene's, or a random-walk medium's
// own; trace_closest below takes the scene's).
RT_HD TravState trav_start(uint32_t root_ref, float tmax) {
    TravState st;
    st.h.t = tmax;
    st.h.b1 = 0.f;
    st.h.b2 = 0.f;
    st.h.prim = -1;
    st.ref = root_ref;
    st.sp = 0;
    return st;
}

template <bool COUNT, bool CULL = true, int OCT = -1, bool EARLY = false>
__device__ __forceinline__ Hit trace_closest_tree(const PackedNode* tree_nodes, const float4* tree_tris,
                                                  uint32_t root_ref, const RayInv& r, float tmin, float tmax,
                                                  uint32_t* stk, int stride, TravCount& cnt, float t_exit = 0.f) {
    TravState st = trav_start(root_ref, tmax);
    trace_closest_run<COUNT, CULL, OCT, EARLY>(tree_nodes, tree_tris, r, tmin, stk, stride, cnt, st, t_exit);
    return st.h;
}

template <bool COUNT, bool CULL = true, int OCT = -1, bool EARLY = false>
__device__ __forceinline__ Hit trace_closest(const DevScene& sc, const RayInv& r, float tmin, float tmax,
                                             uint32_t* stk, int stride, TravCount& cnt, float t_exit = 0.f) {
    return trace_closest_tree<COUNT, CULL, OCT, EARLY>(sc.nodes, sc.tri_isect, sc.root_ref, r, tmin, tmax, stk,
                                                       stride, cnt, t_exit);
}

// Any-hit test of the open segment (tmin, tmax) along r: true at the first
// triangle met from its back (det < 0) -- the faces that a ray in the
// opposite direction, which culls back faces, meets from the front -- with t
// inside the segment.  The nearer child is still visited first, and each
// iteration is one node or one triangle as in trace_closest; the first hit
// ends the search, so the result does not depend on the visiting order.
template <bool COUNT>
__device__ __forceinline__ bool trace_occluded(const DevScene& sc, const RayInv& r, float tmin, float tmax,
                                               uint32_t* stk_generic, int stride, TravCount& cnt) {
    LdsU32* stk = (LdsU32*)stk_generic;
    GlobalV4* nodes = (GlobalV4*)sc.nodes;
    GlobalV4* tris = (GlobalV4*)sc.tri_isect;
    uint32_t ref = sc.root_ref;
    bool leaf = (ref & kLeafBit) != 0;
    uint32_t tn = (ref & ~kLeafBit) >> 3, te = tn + (ref & 7u);
    int sp = 0;
    while (true) {
        uint32_t next;
        if (leaf) {
            if (tn < te) {
                GlobalV4* tp = tris + 3 * (size_t)tn;
                const float4 a = as_f4(tp[0]), e1f = as_f4(tp[1]), e2f = as_f4(tp[2]);
                if (COUNT) cnt.tris++;
                const vec3 e1 = v3(e1f.x, e1f.y, e1f.z), e2 = v3(e2f.x, e2f.y, e2f.z);
                const vec3 s1 = cross(r.d, e2);
                const float det = dot(e1, s1);
                if (det < 0.f) {
                    const float inv = 1.f / det;
                    const vec3 so = r.o - v3(a.x, a.y, a.z);
                    const float b1 = inv * dot(s1, so);
                    const vec3 s2 = cross(so, e1);
                    const float b2 = inv * dot(s2, r.d);
                    const float t = inv * dot(s2, e2);
                    if (b1 >= 0.f && b2 >= 0.f && b1 + b2 <= 1.f && t > tmin && t < tmax) return true;
                }
                ++tn;
            }
            if (tn < te) continue;
            if (sp == 0) break;
            next = stk[--sp * stride];
        } else {
            GlobalV4* np = nodes + 4 * (size_t)ref;
            float4 w0 = as_f4(np[0]), w1 = as_f4(np[1]), w2 = as_f4(np[2]);
            const NativeF4 w3f = np[3];
            const uint32_t rl = (uint32_t)f_as_i(w3f.x), rr = (uint32_t)f_as_i(w3f.y);
            if (COUNT) cnt.nodes++;
            float tr, tl;
            bool hr = box_hit(r, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w, tmin, tmax, tr);
            bool hl = box_hit(r, w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, tmin, tmax, tl);
            if (hr && hl) {
                stk[sp++ * stride] = tr < tl ? rl : rr;
                next = tr < tl ? rr : rl;
            } else if (hr || hl) {
                next = hr ? rr : rl;
            } else {
                if (sp == 0) break;
                next = stk[--sp * stride];
            }
        }
        leaf = (next & kLeafBit) != 0;
        ref = next;
        tn = (next & ~kLeafBit) >> 3;
        te = tn + (next & 7u);
    }
    return false;
}

// Octant of a ray as box_hit uses it (bit k: inverse direction negative on axis k).
RT_HD int ray_octant(const RayInv& r) { return (r.nx ? 1 : 0) | (r.ny ? 2 : 0) | (r.nz ? 4 : 0); }

// trace_closest_run with the slab selects resolved at compile time when every
// active lane of the wave has the same octant (after the ray sort most waves
// do, and so do most camera rays); otherwise the general loop.  Same hits
// either way.
template <bool COUNT, bool CULL = true, bool EARLY = false, int SUSPEND = 0>
__device__ __forceinline__ bool trace_closest_run_wave(const DevScene& sc, const RayInv& r, float tmin, uint32_t* stk,
                                                       int stride, TravCount& cnt, TravState& st, float t_exit = 0.f) {
    const int oct = ray_octant(r);
    const int first = __builtin_amdgcn_readfirstlane(oct);
#define RT_RUN(O) \
    trace_closest_run<COUNT, CULL, O, EARLY, SUSPEND>(sc.nodes, sc.tri_isect, r, tmin, stk, stride, cnt, st, t_exit)
    if (__ballot(oct != first) == 0ull) {
        switch (first) {
            case 0: return RT_RUN(0);
            case 1: return RT_RUN(1);
            case 2: return RT_RUN(2);
            case 3: return RT_RUN(3);
            case 4: return RT_RUN(4);
            case 5: return RT_RUN(5);
            case 6: return RT_RUN(6);
            default: return RT_RUN(7);
        }
    }
    return RT_RUN(-1);
#undef RT_RUN
}

template <bool COUNT, bool CULL = true, bool EARLY = false>
__device__ __forceinline__ Hit trace_closest_wave(const DevScene& sc, const RayInv& r, float tmin, float tmax,
                                                  uint32_t* stk, int stride, TravCount& cnt, float t_exit = 0.f) {
    TravState st = trav_start(sc.root_ref, tmax);
    trace_closest_run_wave<COUNT, CULL, EARLY>(sc, r, tmin, stk, stride, cnt, st, t_exit);
    return st.h;
}

// Shading data at a hit, reconstructed from barycentrics the way the
// reference's triangle test records it (Src/Core/triangle.cu:121-131).
struct SurfaceHit {
    vec3 p;       // b0 v0 + b1 v1 + b2 v2
    vec3 ns;      // normalised interpolated shading normal
    vec3 ng;      // stored geometric normal
    float u, v;
    int mesh, material, light;
};

RT_HD SurfaceHit surface_at(const float4* attr, int prim, float b1, float b2) {
    const float4* a = attr + kAttrStride * (size_t)prim;
    float4 a0 = ldg4(a), a1 = ldg4(a + 1), a2 = ldg4(a + 2), a3 = ldg4(a + 3), a4 = ldg4(a + 4), a5 = ldg4(a + 5),
           a6 = ldg4(a + 6), a7 = ldg4(a + 7);
    float b0 = (1.f - b1) - b2;
    SurfaceHit s;
    vec3 v0 = v3(a0.x, a0.y, a0.z), v1 = v3(a1.x, a1.y, a1.z), v2 = v3(a2.x, a2.y, a2.z);
    s.p = (b0 * v0 + b1 * v1) + b2 * v2;
    vec3 n0 = v3(a3.x, a3.y, a3.z), n1 = v3(a4.x, a4.y, a4.z), n2 = v3(a5.x, a5.y, a5.z);
    s.ns = normalize((b0 * n0 + b1 * n1) + b2 * n2);
    s.ng = v3(a1.w, a2.w, a3.w);
    s.u = (b0 * a4.w + b1 * a6.x) + b2 * a6.z;
    s.v = (b0 * a5.w + b1 * a6.y) + b2 * a6.w;
    s.mesh = f_as_i(a7.x);
    s.material = f_as_i(a7.y);
    s.light = f_as_i(a7.z);
    return s;
}

RT_HD int shade_class_of(const float4* attr, int prim) {
    return prim < 0 ? (int)kClsTerminal : f_as_i(ldg4(attr + kAttrStride * (size_t)prim + 7).w);
}

RT_HD vec3 position_at(const float4* attr, int prim, float b1, float b2) {
    const float4* a = attr + kAttrStride * (size_t)prim;
    float4 a0 = ldg4(a), a1 = ldg4(a + 1), a2 = ldg4(a + 2);
    float b0 = (1.f - b1) - b2;
    vec3 v0 = v3(a0.x, a0.y, a0.z), v1 = v3(a1.x, a1.y, a1.z), v2 = v3(a2.x, a2.y, a2.z);
    return (b0 * v0 + b1 * v1) + b2 * v2;
}

}  // namespace rt
