// GPU SAH BVH build.  Produces exactly the host build's result (bvh.cpp,
// restating Src/Core/bvh.cu:44-206): the same splits, the same triangle
// order and the same DFS node numbering, so traversal is unchanged.
//
// The tree is built level by level: one workgroup per node of a level (one
// wave, or 16 waves for a node of more than 2048 triangles), and for a node
// of more than 65536 triangles chunks of 16384 positions spread over many
// workgroups, which combine through global atomics and per-node scans: node bounds, centroid bounds, the 12-bucket SAH histogram, the split
// choice and the partition of the node's range into the other of two
// position-ordered buffers.  Split decisions are per-node FP32 arithmetic in
// the host's operation order; the pieces that depend on the order of the
// sequential loops are reproduced exactly:
//  - node bounds keep the first triangle (in range order) among equal
//    extremes, which fixes the sign of a zero bound: each bound is reduced as
//    a 64-bit key (order-preserving value, then position);
//  - the partition is libstdc++'s two-pointer std::partition (bvh.cu:159-165):
//    the k-th element of the left zone [start, start + L) that fails the
//    predicate trades places with the k-th element from the right of the right
//    zone that passes it, everything else stays;
//  - leaves hold their ranges in DFS order, so a leaf's first ordered
//    triangle is its range start.
// Centroid and bucket boxes are reduced in any order: they only feed the
// bucket index and the costs, where a zero's sign cannot change a result.
// Node numbering (DFS pre-order, interior-node compaction) is computed on the
// host from the node records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <vector>

#include "bvh.h"
#include "engine.h"

namespace rt {
namespace {

constexpr float kFloatMax = (float)3.40282e+038;
constexpr int kBuckets = 12;
constexpr int kMaxLeaf = 4;
constexpr int kWave = 64;
constexpr int kMediumNode = 2048;   // more triangles: a 16-wave workgroup per node
constexpr int kHugeNode = 65536;    // more triangles: chunks of kChunk over many workgroups
constexpr int kChunk = 16384;

__device__ inline float dmin(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ inline float dmax(float a, float b) { return (a < b) ? b : a; }  // std::max

// Order-preserving key of a float (-0 and +0 compare equal, as in the sequential loop).
__device__ inline uint32_t ord(float v) {
    uint32_t b = __float_as_uint(v);
    if (v == 0.0f) b = 0u;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float unord(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

__device__ inline uint64_t wave_min_u64(uint64_t v) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, kWave);
        v = w < v ? w : v;
    }
    return v;
}
__device__ inline uint64_t wave_max_u64(uint64_t v) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}
__device__ inline float wave_min_f(float v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline float wave_max_f(float v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ inline int lanes_below(uint64_t mask) {
    return __popcll(mask & ((1ull << __lane_id()) - 1ull));
}

// Per-position triangle record, SoA: (lo.xyz, centroid.x), (hi.xyz, centroid.y), (centroid.z, index bits).
struct Elems {
    float4* lo_cx;
    float4* hi_cy;
    float2* cz_id;
};

struct Nodes {
    int4* range;     // (start, end, depth, flags: 1 leaf, 2 fallback split)
    int2* kids;      // creation ids of the children
    float4* box_lo;  // node bounds
    float4* box_hi;
};

__global__ void k_bvh_init(const float* __restrict__ pos, int n, Elems e) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float* p = pos + 9 * (size_t)t;
    // Bounds3D(v0, v1, v2) = Min(v0, Min(v1, v2)) (Src/Core/bounds.h:22-26).
    float lx = dmin(p[0], dmin(p[3], p[6])), ly = dmin(p[1], dmin(p[4], p[7])), lz = dmin(p[2], dmin(p[5], p[8]));
    float hx = dmax(p[0], dmax(p[3], p[6])), hy = dmax(p[1], dmax(p[4], p[7])), hz = dmax(p[2], dmax(p[5], p[8]));
    e.lo_cx[t] = make_float4(lx, ly, lz, (lx + hx) * 0.5f);
    e.hi_cy[t] = make_float4(hx, hy, hz, (ly + hy) * 0.5f);
    e.cz_id[t] = make_float2((lz + hz) * 0.5f, __int_as_float(t));
}

// Bucket of a centroid (Bounds3D::Offset, Src/Core/bounds.h:89-97; bvh.cu:100-108).
__device__ inline int bucket_of(float c, float lo, float hi) {
    float o = c - lo;
    float ext = hi - lo;
    if (ext > 0) o /= ext;
    int b = (int)((float)kBuckets * o);
    return b == kBuckets ? kBuckets - 1 : b;
}

// Costs of splitting after bucket i (bvh.cu:117-144): the right side sums
// buckets i+1..10 (bucket 11 is left out) and the search starts at i = 1.
// Bucket boxes are order keys; box = node bounds (lo.xyz, hi.xyz).
__device__ inline float surface_area(const float* l, const float* h) {
    if (l[0] > h[0] || l[1] > h[1] || l[2] > h[2]) return 0.f;
    float e0 = h[0] - l[0], e1 = h[1] - l[1], e2 = h[2] - l[2];
    return 2.f * ((e0 * e1 + e1 * e2) + e2 * e0);
}
__device__ inline int choose_split(const uint32_t* cnt, const uint32_t (*blo)[3], const uint32_t (*bhi)[3],
                                   const float* box) {
    const float sa_node = surface_area(box, box + 3);
    float min_cost = kFloatMax;
    int best = -1;
    for (int i = 1; i < kBuckets - 1; ++i) {
        float b0l[3] = {kFloatMax, kFloatMax, kFloatMax}, b0h[3] = {-kFloatMax, -kFloatMax, -kFloatMax};
        float b1l[3] = {kFloatMax, kFloatMax, kFloatMax}, b1h[3] = {-kFloatMax, -kFloatMax, -kFloatMax};
        int c0 = 0, c1 = 0;
        for (int j = 0; j <= i; ++j) {
            for (int k = 0; k < 3; ++k) {
                b0l[k] = dmin(unord(blo[j][k]), b0l[k]);
                b0h[k] = dmax(unord(bhi[j][k]), b0h[k]);
            }
            c0 += (int)cnt[j];
        }
        for (int j = i + 1; j < kBuckets - 1; ++j) {
            for (int k = 0; k < 3; ++k) {
                b1l[k] = dmin(unord(blo[j][k]), b1l[k]);
                b1h[k] = dmax(unord(bhi[j][k]), b1h[k]);
            }
            c1 += (int)cnt[j];
        }
        float cost = .125f + ((float)c0 * surface_area(b0l, b0h) + (float)c1 * surface_area(b1l, b1h)) / sa_node;
        if (cost < min_cost) {
            min_cost = cost;
            best = i;
        }
    }
    return best;
}

// Block-wide helpers for a workgroup of WAVES waves (results in every thread).
template <int WAVES>
struct BlockOps {
    uint64_t* u64;  // WAVES entries of LDS
    float* f32;     // WAVES entries
    int* i32;       // WAVES entries
    __device__ uint64_t min_u64(uint64_t v) {
        v = wave_min_u64(v);
        if (WAVES == 1) return v;
        __syncthreads();
        if (__lane_id() == 0) u64[threadIdx.x / kWave] = v;
        __syncthreads();
        uint64_t r = u64[0];
        for (int w = 1; w < WAVES; ++w) r = u64[w] < r ? u64[w] : r;
        return r;
    }
    __device__ uint64_t max_u64(uint64_t v) {
        v = wave_max_u64(v);
        if (WAVES == 1) return v;
        __syncthreads();
        if (__lane_id() == 0) u64[threadIdx.x / kWave] = v;
        __syncthreads();
        uint64_t r = u64[0];
        for (int w = 1; w < WAVES; ++w) r = u64[w] > r ? u64[w] : r;
        return r;
    }
    __device__ float min_f(float v) {
        v = wave_min_f(v);
        if (WAVES == 1) return v;
        __syncthreads();
        if (__lane_id() == 0) f32[threadIdx.x / kWave] = v;
        __syncthreads();
        float r = f32[0];
        for (int w = 1; w < WAVES; ++w) r = dmin(r, f32[w]);
        return r;
    }
    __device__ float max_f(float v) {
        v = wave_max_f(v);
        if (WAVES == 1) return v;
        __syncthreads();
        if (__lane_id() == 0) f32[threadIdx.x / kWave] = v;
        __syncthreads();
        float r = f32[0];
        for (int w = 1; w < WAVES; ++w) r = dmax(r, f32[w]);
        return r;
    }
    // Rank of this thread's flag among the flags of threads before it (thread
    // order = position order of the step) and the step's total.
    __device__ int rank(bool flag, int& total) {
        uint64_t mask = __ballot(flag);
        int below = lanes_below(mask);
        if (WAVES == 1) {
            total = __popcll(mask);
            return below;
        }
        __syncthreads();
        if (__lane_id() == 0) i32[threadIdx.x / kWave] = __popcll(mask);
        __syncthreads();
        int before = 0, sum = 0;
        for (int w = 0; w < WAVES; ++w) {
            int c = i32[w];
            if (w < (int)(threadIdx.x / kWave)) before += c;
            sum += c;
        }
        total = sum;
        return before + below;
    }
};

// A node of at most 64 triangles in one wave, one triangle per lane: the
// same steps as below with every element read once and kept in registers,
// ranks from ballots and the two-pointer pairing through LDS.
__device__ void node_in_registers(const Elems& cur, const Elems& nxt, int32_t* __restrict__ order, const Nodes& nodes,
                                  int node, int4 rg, int node_cap, int* __restrict__ counters, uint32_t* s_cnt,
                                  uint32_t (*s_lo)[3], uint32_t (*s_hi)[3]) {
    __shared__ int s_fail[kWave], s_pass[kWave];
    const int lane = __lane_id();
    const int start = rg.x, end = rg.y, n = end - start;
    const int p = start + lane;
    const bool valid = lane < n;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
    float2 c = make_float2(0.f, 0.f);
    if (valid) {
        a = cur.lo_cx[p];
        b = cur.hi_cy[p];
        c = cur.cz_id[p];
    }
    const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z}, ce[3] = {a.w, b.w, c.x};
    float box[6], cb_lo[3], cb_hi[3];
    for (int k = 0; k < 3; ++k) {
        uint64_t kl = valid ? ((uint64_t)ord(lo[k]) << 32 | (uint32_t)p) : ~0ull;
        uint64_t kh = valid ? ((uint64_t)ord(hi[k]) << 32 | (0xffffffffu - (uint32_t)p)) : 0ull;
        int wl = (int)(uint32_t)wave_min_u64(kl) - start, wh = (int)(0xffffffffu - (uint32_t)wave_max_u64(kh)) - start;
        box[k] = __shfl(lo[k], wl, kWave);
        box[3 + k] = __shfl(hi[k], wh, kWave);
        cb_lo[k] = wave_min_f(valid ? ce[k] : kFloatMax);
        cb_hi[k] = wave_max_f(valid ? ce[k] : -kFloatMax);
    }
    if (lane == 0) {
        nodes.box_lo[node] = make_float4(box[0], box[1], box[2], 0.f);
        nodes.box_hi[node] = make_float4(box[3], box[4], box[5], 0.f);
    }
    if (n <= kMaxLeaf) {
        if (valid) order[p] = __float_as_int(c.y);
        if (lane == 0) nodes.range[node].w = 1;
        return;
    }
    const float ex = cb_hi[0] - cb_lo[0], ey = cb_hi[1] - cb_lo[1], ez = cb_hi[2] - cb_lo[2];
    int axis = 0;
    float m = ex;
    if (ey > m) { m = ey; axis = 1; }
    if (ez > m) axis = 2;
    const int bk = bucket_of(ce[axis], cb_lo[axis], cb_hi[axis]);
    if (lane < kBuckets) {
        s_cnt[lane] = 0u;
        for (int k = 0; k < 3; ++k) {
            s_lo[lane][k] = ord(kFloatMax);
            s_hi[lane][k] = ord(-kFloatMax);
        }
    }
    __syncthreads();
    if (valid) {
        atomicAdd(&s_cnt[bk], 1u);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&s_lo[bk][k], ord(lo[k]));
            atomicMax(&s_hi[bk][k], ord(hi[k]));
        }
    }
    __syncthreads();
    const int best = choose_split(s_cnt, s_lo, s_hi, box);
    int L = 0;
    for (int j = 0; j <= best; ++j) L += (int)s_cnt[j];
    const int zone = start + L;
    const bool fallback = best < 0 || zone <= start || zone >= end;
    const int mid = fallback ? (start + end) / 2 : zone;
    int dst = p;
    if (!fallback) {
        const bool pass = valid && bk <= best;
        const bool fl = valid && p < zone && !pass, pr = valid && p >= zone && pass;
        const uint64_t mf = __ballot(fl), mp = __ballot(pr);
        const int kf = lanes_below(mf);
        const int kp = __popcll(mp & ~((2ull << lane) - 1ull));  // passing lanes above this one
        if (fl) s_fail[kf] = p;
        if (pr) s_pass[kp] = p;
        __syncthreads();
        if (fl) dst = s_pass[kf];
        if (pr) dst = s_fail[kp];
    }
    if (valid) {
        nxt.lo_cx[dst] = a;
        nxt.hi_cy[dst] = b;
        nxt.cz_id[dst] = c;
    }
    if (lane == 0) {
        int id = atomicAdd(&counters[0], 2);
        if (id + 2 > node_cap) {
            atomicOr(&counters[1], 2);
            return;
        }
        nodes.range[id] = make_int4(start, mid, rg.z + 1, 0);
        nodes.range[id + 1] = make_int4(mid, end, rg.z + 1, 0);
        nodes.kids[node] = make_int2(id, id + 1);
        nodes.range[node].w = fallback ? 2 : 0;
    }
}

// One node per workgroup of WAVES waves; `list` names the nodes of this launch.
template <int WAVES>
__global__ __launch_bounds__(kWave * WAVES) void k_bvh_node(Elems cur, Elems nxt, int32_t* __restrict__ order,
                                                            int32_t* __restrict__ scratch, Nodes nodes,
                                                            const int32_t* __restrict__ list, int node_cap,
                                                            int* __restrict__ counters) {
    constexpr int kThreads = kWave * WAVES;
    __shared__ uint32_t s_cnt[kBuckets];
    __shared__ uint32_t s_lo[kBuckets][3], s_hi[kBuckets][3];
    __shared__ uint64_t s_u64[WAVES];
    __shared__ float s_f32[WAVES];
    __shared__ int s_i32[WAVES];
    BlockOps<WAVES> ops{s_u64, s_f32, s_i32};
    const int node = list[blockIdx.x];
    const int t = threadIdx.x;
    const int4 rg = nodes.range[node];
    const int start = rg.x, end = rg.y, n = end - start;
    if (WAVES == 1 && n <= kWave) {
        node_in_registers(cur, nxt, order, nodes, node, rg, node_cap, counters, s_cnt, s_lo, s_hi);
        return;
    }

    // Node bounds: first position among equal extremes.
    uint64_t klo[3], khi[3];
    float clo[3], chi[3];
    for (int k = 0; k < 3; ++k) {
        klo[k] = ~0ull;
        khi[k] = 0ull;
        clo[k] = kFloatMax;
        chi[k] = -kFloatMax;
    }
    for (int base = start; base < end; base += kThreads) {
        int p = base + t;
        if (p < end) {
            float4 a = cur.lo_cx[p], b = cur.hi_cy[p];
            float2 c = cur.cz_id[p];
            const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z}, ce[3] = {a.w, b.w, c.x};
            for (int k = 0; k < 3; ++k) {
                uint64_t kl = (uint64_t)ord(lo[k]) << 32 | (uint32_t)p;
                uint64_t kh = (uint64_t)ord(hi[k]) << 32 | (0xffffffffu - (uint32_t)p);
                klo[k] = kl < klo[k] ? kl : klo[k];
                khi[k] = kh > khi[k] ? kh : khi[k];
                clo[k] = dmin(clo[k], ce[k]);
                chi[k] = dmax(chi[k], ce[k]);
            }
        }
    }
    float box[6], cb_lo[3], cb_hi[3];
    for (int k = 0; k < 3; ++k) {
        uint64_t ml = ops.min_u64(klo[k]), mh = ops.max_u64(khi[k]);
        int pl = (int)(uint32_t)ml, ph = (int)(0xffffffffu - (uint32_t)mh);
        float4 va = cur.lo_cx[pl], vb = cur.hi_cy[ph];
        box[k] = k == 0 ? va.x : (k == 1 ? va.y : va.z);
        box[3 + k] = k == 0 ? vb.x : (k == 1 ? vb.y : vb.z);
        cb_lo[k] = ops.min_f(clo[k]);
        cb_hi[k] = ops.max_f(chi[k]);
    }
    if (t == 0) {
        nodes.box_lo[node] = make_float4(box[0], box[1], box[2], 0.f);
        nodes.box_hi[node] = make_float4(box[3], box[4], box[5], 0.f);
    }

    if (n <= kMaxLeaf) {
        for (int p = start + t; p < end; p += kThreads) order[p] = __float_as_int(cur.cz_id[p].y);
        if (t == 0) nodes.range[node].w = 1;
        return;
    }

    // Split axis: first index of the largest centroid extent (Eigen maxCoeff, strict '>').
    const float ex = cb_hi[0] - cb_lo[0], ey = cb_hi[1] - cb_lo[1], ez = cb_hi[2] - cb_lo[2];
    int axis = 0;
    float m = ex;
    if (ey > m) { m = ey; axis = 1; }
    if (ez > m) axis = 2;
    const float ax_lo = cb_lo[axis], ax_hi = cb_hi[axis];

    if (t < kBuckets) {
        s_cnt[t] = 0u;
        for (int k = 0; k < 3; ++k) {
            s_lo[t][k] = ord(kFloatMax);
            s_hi[t][k] = ord(-kFloatMax);
        }
    }
    __syncthreads();
    for (int base = start; base < end; base += kThreads) {
        int p = base + t;
        if (p < end) {
            float4 a = cur.lo_cx[p], b = cur.hi_cy[p];
            float c = axis == 0 ? a.w : (axis == 1 ? b.w : cur.cz_id[p].x);
            int bk = bucket_of(c, ax_lo, ax_hi);
            atomicAdd(&s_cnt[bk], 1u);
            atomicMin(&s_lo[bk][0], ord(a.x));
            atomicMin(&s_lo[bk][1], ord(a.y));
            atomicMin(&s_lo[bk][2], ord(a.z));
            atomicMax(&s_hi[bk][0], ord(b.x));
            atomicMax(&s_hi[bk][1], ord(b.y));
            atomicMax(&s_hi[bk][2], ord(b.z));
        }
    }
    __syncthreads();

    const int best = choose_split(s_cnt, s_lo, s_hi, box);
    int L = 0;
    for (int j = 0; j <= best; ++j) L += (int)s_cnt[j];
    const int zone = start + L;
    const bool fallback = best < 0 || zone <= start || zone >= end;
    const int mid = fallback ? (start + end) / 2 : zone;

    if (fallback) {
        // No reordering: the range moves to the next buffer unchanged.
        for (int p = start + t; p < end; p += kThreads) {
            nxt.lo_cx[p] = cur.lo_cx[p];
            nxt.hi_cy[p] = cur.hi_cy[p];
            nxt.cz_id[p] = cur.cz_id[p];
        }
    } else {
        auto passes = [&](int p) {
            float c = axis == 0 ? cur.lo_cx[p].w : (axis == 1 ? cur.hi_cy[p].w : cur.cz_id[p].x);
            return bucket_of(c, ax_lo, ax_hi) <= best;
        };
        // Positions of the left zone's failing elements (left to right) and of
        // the right zone's passing elements (right to left).
        int nfail = 0;
        for (int base = start; base < zone; base += kThreads) {
            int p = base + t;
            bool f = p < zone && !passes(p);
            int tot, r = ops.rank(f, tot);
            if (f) scratch[start + nfail + r] = p;  // < zone: inside the range
            nfail += tot;
        }
        int npass = 0;
        for (int top = end; top > zone; top -= kThreads) {
            int p = top - 1 - t;
            bool s = p >= zone && passes(p);
            int tot, r = ops.rank(s, tot);
            int slot = start + nfail + npass + r;
            if (s && slot < end) scratch[slot] = p;
            npass += tot;
        }
        // The zone holds as many passing elements on the right as failing
        // ones on the left; anything else would be an inconsistent predicate.
        if (npass != nfail) {
            if (t == 0) atomicOr(&counters[1], 1);
            return;
        }
        __threadfence();
        __syncthreads();
        // Scatter: the k-th failing element of the left zone goes to the k-th
        // passing position from the right, and that element to its place.
        int kf = 0;
        for (int base = start; base < zone; base += kThreads) {
            int p = base + t;
            bool valid = p < zone;
            bool f = valid && !passes(p);
            int tot, r = ops.rank(f, tot);
            if (valid) {
                int dst = f ? scratch[start + nfail + kf + r] : p;
                if (dst < start || dst >= end) dst = p;  // unreachable with a consistent predicate
                nxt.lo_cx[dst] = cur.lo_cx[p];
                nxt.hi_cy[dst] = cur.hi_cy[p];
                nxt.cz_id[dst] = cur.cz_id[p];
            }
            kf += tot;
        }
        int ks = 0;
        for (int top = end; top > zone; top -= kThreads) {
            int p = top - 1 - t;
            bool valid = p >= zone;
            bool s = valid && passes(p);
            int tot, r = ops.rank(s, tot);
            if (valid) {
                int dst = s ? scratch[start + ks + r] : p;
                if (dst < start || dst >= end) dst = p;
                nxt.lo_cx[dst] = cur.lo_cx[p];
                nxt.hi_cy[dst] = cur.hi_cy[p];
                nxt.cz_id[dst] = cur.cz_id[p];
            }
            ks += tot;
        }
    }
    if (t == 0) {
        int id = atomicAdd(&counters[0], 2);
        if (id + 2 > node_cap) {
            atomicOr(&counters[1], 2);
            return;
        }
        nodes.range[id] = make_int4(start, mid, rg.z + 1, 0);
        nodes.range[id + 1] = make_int4(mid, end, rg.z + 1, 0);
        nodes.kids[node] = make_int2(id, id + 1);
        nodes.range[node].w = fallback ? 2 : 0;
    }
}

// ---- nodes of more than kHugeNode triangles: many workgroups each
// (chunks of kChunk positions, "tasks"), reductions through global atomics,
// ranks from per-chunk counts scanned per node.
struct HugeAcc {
    unsigned long long klo[3], khi[3];  // node bounds as (order key, position) keys
    uint32_t clo[3], chi[3];            // centroid bounds as order keys
    uint32_t cnt[kBuckets];
    uint32_t blo[kBuckets][3], bhi[kBuckets][3];
    float box[6];
    float ax_lo, ax_hi;
    int axis, zone, mid, fallback, nfail, best;
};

struct HugeTask {
    int slot, first, last, node;  // node's accumulator slot and the chunk [first, last)
};

__global__ void k_huge_init(HugeAcc* __restrict__ acc, int nslots) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nslots) return;
    HugeAcc& a = acc[j];
    for (int k = 0; k < 3; ++k) {
        a.klo[k] = ~0ull;
        a.khi[k] = 0ull;
        a.clo[k] = ord(kFloatMax);
        a.chi[k] = ord(-kFloatMax);
    }
    for (int b = 0; b < kBuckets; ++b) {
        a.cnt[b] = 0u;
        for (int k = 0; k < 3; ++k) {
            a.blo[b][k] = ord(kFloatMax);
            a.bhi[b][k] = ord(-kFloatMax);
        }
    }
}

constexpr int kHugeWaves = 16;
constexpr int kHugeThreads = kWave * kHugeWaves;

__global__ __launch_bounds__(kHugeThreads) void k_huge_bounds(Elems cur, const HugeTask* __restrict__ tasks,
                                                              HugeAcc* __restrict__ acc) {
    __shared__ uint64_t s_u64[kHugeWaves];
    __shared__ float s_f32[kHugeWaves];
    __shared__ int s_i32[kHugeWaves];
    BlockOps<kHugeWaves> ops{s_u64, s_f32, s_i32};
    const HugeTask tk = tasks[blockIdx.x];
    uint64_t klo[3], khi[3];
    float clo[3], chi[3];
    for (int k = 0; k < 3; ++k) {
        klo[k] = ~0ull;
        khi[k] = 0ull;
        clo[k] = kFloatMax;
        chi[k] = -kFloatMax;
    }
    for (int p = tk.first + (int)threadIdx.x; p < tk.last; p += kHugeThreads) {
        float4 a = cur.lo_cx[p], b = cur.hi_cy[p];
        float2 c = cur.cz_id[p];
        const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z}, ce[3] = {a.w, b.w, c.x};
        for (int k = 0; k < 3; ++k) {
            uint64_t kl = (uint64_t)ord(lo[k]) << 32 | (uint32_t)p;
            uint64_t kh = (uint64_t)ord(hi[k]) << 32 | (0xffffffffu - (uint32_t)p);
            klo[k] = kl < klo[k] ? kl : klo[k];
            khi[k] = kh > khi[k] ? kh : khi[k];
            clo[k] = dmin(clo[k], ce[k]);
            chi[k] = dmax(chi[k], ce[k]);
        }
    }
    HugeAcc& A = acc[tk.slot];
    for (int k = 0; k < 3; ++k) {
        uint64_t ml = ops.min_u64(klo[k]), mh = ops.max_u64(khi[k]);
        float fl = ops.min_f(clo[k]), fh = ops.max_f(chi[k]);
        if (threadIdx.x == 0) {
            atomicMin(&A.klo[k], (unsigned long long)ml);
            atomicMax(&A.khi[k], (unsigned long long)mh);
            atomicMin(&A.clo[k], ord(fl));
            atomicMax(&A.chi[k], ord(fh));
        }
    }
}

// Per node: bounds from the winning positions, split axis from the centroid bounds.
__global__ void k_huge_axis(Elems cur, const int4* __restrict__ huge, int nslots, HugeAcc* __restrict__ acc,
                            Nodes nodes) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nslots) return;
    HugeAcc& A = acc[j];
    const int node = huge[j].x;
    float box[6];
    for (int k = 0; k < 3; ++k) {
        int pl = (int)(uint32_t)A.klo[k], ph = (int)(0xffffffffu - (uint32_t)A.khi[k]);
        float4 va = cur.lo_cx[pl], vb = cur.hi_cy[ph];
        box[k] = k == 0 ? va.x : (k == 1 ? va.y : va.z);
        box[3 + k] = k == 0 ? vb.x : (k == 1 ? vb.y : vb.z);
    }
    for (int k = 0; k < 6; ++k) A.box[k] = box[k];
    nodes.box_lo[node] = make_float4(box[0], box[1], box[2], 0.f);
    nodes.box_hi[node] = make_float4(box[3], box[4], box[5], 0.f);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = unord(A.clo[k]);
        hi[k] = unord(A.chi[k]);
    }
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    int axis = 0;
    float m = ex;
    if (ey > m) { m = ey; axis = 1; }
    if (ez > m) axis = 2;
    A.axis = axis;
    A.ax_lo = lo[axis];
    A.ax_hi = hi[axis];
}

__global__ __launch_bounds__(kHugeThreads) void k_huge_buckets(Elems cur, const HugeTask* __restrict__ tasks,
                                                               HugeAcc* __restrict__ acc) {
    __shared__ uint32_t s_cnt[kBuckets];
    __shared__ uint32_t s_lo[kBuckets][3], s_hi[kBuckets][3];
    const HugeTask tk = tasks[blockIdx.x];
    HugeAcc& A = acc[tk.slot];
    const int axis = A.axis;
    const float ax_lo = A.ax_lo, ax_hi = A.ax_hi;
    const int t = threadIdx.x;
    if (t < kBuckets) {
        s_cnt[t] = 0u;
        for (int k = 0; k < 3; ++k) {
            s_lo[t][k] = ord(kFloatMax);
            s_hi[t][k] = ord(-kFloatMax);
        }
    }
    __syncthreads();
    for (int p = tk.first + t; p < tk.last; p += kHugeThreads) {
        float4 a = cur.lo_cx[p], b = cur.hi_cy[p];
        float c = axis == 0 ? a.w : (axis == 1 ? b.w : cur.cz_id[p].x);
        int bk = bucket_of(c, ax_lo, ax_hi);
        atomicAdd(&s_cnt[bk], 1u);
        atomicMin(&s_lo[bk][0], ord(a.x));
        atomicMin(&s_lo[bk][1], ord(a.y));
        atomicMin(&s_lo[bk][2], ord(a.z));
        atomicMax(&s_hi[bk][0], ord(b.x));
        atomicMax(&s_hi[bk][1], ord(b.y));
        atomicMax(&s_hi[bk][2], ord(b.z));
    }
    __syncthreads();
    if (t < kBuckets && s_cnt[t]) {
        atomicAdd(&A.cnt[t], s_cnt[t]);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&A.blo[t][k], s_lo[t][k]);
            atomicMax(&A.bhi[t][k], s_hi[t][k]);
        }
    }
}

// Per node: split choice, children.
__global__ void k_huge_split(const int4* __restrict__ huge, int nslots, HugeAcc* __restrict__ acc, Nodes nodes,
                             int node_cap, int* __restrict__ counters) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nslots) return;
    HugeAcc& A = acc[j];
    const int4 h = huge[j];
    const int node = h.x, start = h.y, end = h.z;
    const int best = choose_split(A.cnt, A.blo, A.bhi, A.box);
    int L = 0;
    for (int b = 0; b <= best; ++b) L += (int)A.cnt[b];
    const int zone = start + L;
    const bool fallback = best < 0 || zone <= start || zone >= end;
    const int mid = fallback ? (start + end) / 2 : zone;
    A.best = best;
    A.zone = zone;
    A.fallback = fallback ? 1 : 0;
    A.mid = mid;
    const int depth = nodes.range[node].z;
    int id = atomicAdd(&counters[0], 2);
    if (id + 2 > node_cap) {
        atomicOr(&counters[1], 2);
        A.fallback = -1;  // nothing more is done for this node
        return;
    }
    nodes.range[id] = make_int4(start, mid, depth + 1, 0);
    nodes.range[id + 1] = make_int4(mid, end, depth + 1, 0);
    nodes.kids[node] = make_int2(id, id + 1);
    nodes.range[node].w = fallback ? 2 : 0;
}

__device__ inline bool huge_passes(const Elems& cur, const HugeAcc& A, int p) {
    float c = A.axis == 0 ? cur.lo_cx[p].w : (A.axis == 1 ? cur.hi_cy[p].w : cur.cz_id[p].x);
    return bucket_of(c, A.ax_lo, A.ax_hi) <= A.best;
}

// Per chunk: failing elements left of the zone boundary, passing elements right of it.
__global__ __launch_bounds__(kHugeThreads) void k_huge_count(Elems cur, const HugeTask* __restrict__ tasks,
                                                             const HugeAcc* __restrict__ acc,
                                                             int2* __restrict__ task_cnt) {
    __shared__ int s_i32[kHugeWaves];
    const HugeTask tk = tasks[blockIdx.x];
    const HugeAcc& A = acc[tk.slot];
    if (A.fallback) return;
    int fl = 0, pr = 0;
    for (int p = tk.first + (int)threadIdx.x; p < tk.last; p += kHugeThreads) {
        bool ps = huge_passes(cur, A, p);
        if (p < A.zone) fl += !ps;
        else pr += ps;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        fl += __shfl_xor(fl, o, kWave);
        pr += __shfl_xor(pr, o, kWave);
    }
    if (__lane_id() == 0) s_i32[threadIdx.x / kWave] = fl;
    __syncthreads();
    int tfl = 0;
    for (int w = 0; w < kHugeWaves; ++w) tfl += s_i32[w];
    __syncthreads();
    if (__lane_id() == 0) s_i32[threadIdx.x / kWave] = pr;
    __syncthreads();
    int tpr = 0;
    for (int w = 0; w < kHugeWaves; ++w) tpr += s_i32[w];
    if (threadIdx.x == 0) task_cnt[blockIdx.x] = make_int2(tfl, tpr);
}

// Per node: offsets of each chunk's ranks (failing from the left, passing from the right).
__global__ void k_huge_scan(const int* __restrict__ task_first, int nslots, HugeAcc* __restrict__ acc,
                            const int2* __restrict__ task_cnt, int2* __restrict__ task_off,
                            int* __restrict__ counters) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nslots) return;
    HugeAcc& A = acc[j];
    if (A.fallback) return;
    const int t0 = task_first[j], t1 = task_first[j + 1];
    int fl = 0, pr = 0;
    for (int t = t0; t < t1; ++t) {
        int2 c = task_cnt[t];
        task_off[t] = make_int2(fl, pr);  // .y: passing elements in earlier chunks, for now
        fl += c.x;
        pr += c.y;
    }
    for (int t = t0; t < t1; ++t) {
        int2 o = task_off[t];
        task_off[t] = make_int2(o.x, pr - o.y - task_cnt[t].y);  // passing elements in later chunks
    }
    A.nfail = fl;
    if (fl != pr) {
        atomicOr(&counters[1], 1);
        A.fallback = -1;
    }
}

// Per chunk: record the positions of the ranked elements (PLACE) or move every
// element to its destination in the next buffer (!PLACE).
template <bool PLACE>
__global__ __launch_bounds__(kHugeThreads) void k_huge_move(Elems cur, Elems nxt, const int4* __restrict__ huge,
                                                            const HugeTask* __restrict__ tasks,
                                                            const HugeAcc* __restrict__ acc,
                                                            const int2* __restrict__ task_off,
                                                            int32_t* __restrict__ scratch) {
    __shared__ uint64_t s_u64[kHugeWaves];
    __shared__ float s_f32[kHugeWaves];
    __shared__ int s_i32[kHugeWaves];
    BlockOps<kHugeWaves> ops{s_u64, s_f32, s_i32};
    const HugeTask tk = tasks[blockIdx.x];
    const HugeAcc& A = acc[tk.slot];
    const int t = threadIdx.x;
    if (A.fallback < 0) return;
    if (A.fallback) {
        if (!PLACE)
            for (int p = tk.first + t; p < tk.last; p += kHugeThreads) {
                nxt.lo_cx[p] = cur.lo_cx[p];
                nxt.hi_cy[p] = cur.hi_cy[p];
                nxt.cz_id[p] = cur.cz_id[p];
            }
        return;
    }
    const int start = huge[tk.slot].y, end = huge[tk.slot].z, zone = A.zone, nfail = A.nfail;
    const int2 off = task_off[blockIdx.x];
    // Left of the zone boundary, left to right.
    const int lz = min(tk.last, zone);
    int run = 0;
    for (int base = tk.first; base < lz; base += kHugeThreads) {
        int p = base + t;
        bool valid = p < lz;
        bool f = valid && !huge_passes(cur, A, p);
        int tot, r = ops.rank(f, tot);
        int k = off.x + run + r;
        if (PLACE) {
            if (f) scratch[start + k] = p;
        } else if (valid) {
            int dst = f ? scratch[start + nfail + k] : p;
            if (dst < start || dst >= end) dst = p;  // unreachable with consistent counts
            nxt.lo_cx[dst] = cur.lo_cx[p];
            nxt.hi_cy[dst] = cur.hi_cy[p];
            nxt.cz_id[dst] = cur.cz_id[p];
        }
        run += tot;
    }
    // Right of it, right to left.
    const int rz = max(tk.first, zone);
    run = 0;
    for (int top = tk.last; top > rz; top -= kHugeThreads) {
        int p = top - 1 - t;
        bool valid = p >= rz;
        bool s = valid && huge_passes(cur, A, p);
        int tot, r = ops.rank(s, tot);
        int k = off.y + run + r;
        if (PLACE) {
            if (s && start + nfail + k < end) scratch[start + nfail + k] = p;
        } else if (valid) {
            int dst = s ? scratch[start + k] : p;
            if (dst < start || dst >= end) dst = p;
            nxt.lo_cx[dst] = cur.lo_cx[p];
            nxt.hi_cy[dst] = cur.hi_cy[p];
            nxt.cz_id[dst] = cur.cz_id[p];
        }
        run += tot;
    }
}

// Nodes [first, last) of a level by size: up to kMediumNode triangles (one
// wave each), up to kHugeNode (one 16-wave workgroup each), larger (chunk tasks).
__global__ void k_bvh_classify(const int4* __restrict__ range, int first, int last,
                               int32_t* __restrict__ small_list, int32_t* __restrict__ medium_list,
                               int4* __restrict__ huge_list, int* __restrict__ counters) {
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    int4 r = range[i];
    int n = r.y - r.x;
    if (n > kHugeNode) huge_list[atomicAdd(&counters[5], 1)] = make_int4(i, r.x, r.y, 0);
    else if (n > kMediumNode) medium_list[atomicAdd(&counters[3], 1)] = i;
    else small_list[atomicAdd(&counters[2], 1)] = i;
}

// Node numbering from the node records: subtree sizes (nodes, interior
// nodes) bottom-up one level at a time, then DFS pre-order indices and
// interior-node compaction top-down (children are created after their parents).
__global__ void k_bvh_sizes(const int4* __restrict__ range, const int2* __restrict__ kids, int first, int last,
                            int32_t* __restrict__ size, int32_t* __restrict__ icount) {
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    if (range[i].w == 1) {
        size[i] = 1;
        icount[i] = 0;
    } else {
        int2 c = kids[i];
        size[i] = 1 + size[c.x] + size[c.y];
        icount[i] = 1 + icount[c.x] + icount[c.y];
    }
}

__global__ void k_bvh_index(const int4* __restrict__ range, const int2* __restrict__ kids, int first, int last,
                            const int32_t* __restrict__ size, const int32_t* __restrict__ icount,
                            int32_t* __restrict__ dfs, int32_t* __restrict__ comp) {
    int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    if (i == 0) {
        dfs[0] = 0;
        comp[0] = 0;
    }
    if (range[i].w == 1) return;
    int2 c = kids[i];
    int d = i == 0 ? 0 : dfs[i], k = i == 0 ? 0 : comp[i];
    dfs[c.x] = d + 1;
    dfs[c.y] = d + 1 + size[c.x];
    comp[c.x] = k + 1;
    comp[c.y] = k + 1 + icount[c.x];
}

__device__ inline uint32_t child_ref(const int4* range, const int32_t* comp, int c) {
    int4 r = range[c];
    return r.w == 1 ? leaf_ref((uint32_t)r.x, (uint32_t)(r.y - r.x)) : (uint32_t)comp[c];
}

// The reference-layout node array (DFS pre-order) and the packed child-pair lines.
__global__ void k_bvh_emit(Nodes nodes, int nn, const int32_t* __restrict__ dfs, const int32_t* __restrict__ comp,
                           LinearNode* __restrict__ lin, PackedNode* __restrict__ packed, int* __restrict__ counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    int4 r = nodes.range[i];
    float4 lo = nodes.box_lo[i], hi = nodes.box_hi[i];
    LinearNode ln;
    const bool leaf = r.w == 1;
    int2 c = leaf ? make_int2(0, 0) : nodes.kids[i];
    ln.tri_index = leaf ? r.x : -1;
    ln.second = leaf ? r.y - r.x : dfs[c.y];
    ln.lo[0] = lo.x; ln.lo[1] = lo.y; ln.lo[2] = lo.z;
    ln.hi[0] = hi.x; ln.hi[1] = hi.y; ln.hi[2] = hi.z;
    lin[dfs[i]] = ln;
    if (r.w == 2) atomicAdd(&counters[4], 1);
    if (leaf) return;
    float4 ll = nodes.box_lo[c.x], lh = nodes.box_hi[c.x], rl = nodes.box_lo[c.y], rh = nodes.box_hi[c.y];
    PackedNode p;
    p.w0 = make_float4(ll.x, ll.y, ll.z, lh.x);
    p.w1 = make_float4(lh.y, lh.z, rl.x, rl.y);
    p.w2 = make_float4(rl.z, rh.x, rh.y, rh.z);
    p.w3 = make_int4((int)child_ref(nodes.range, comp, c.x), (int)child_ref(nodes.range, comp, c.y), dfs[i], 0);
    packed[comp[i]] = p;
}

// Device buffer allocated and freed in stream order on the build's stream
// (no device-wide synchronisation, unlike hipMalloc / hipFree).
struct StreamBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    explicit StreamBuf(hipStream_t stream) : s(stream) {}
    StreamBuf(const StreamBuf&) = delete;
    StreamBuf& operator=(const StreamBuf&) = delete;
    void alloc(size_t bytes) { HIP_OK(hipMallocAsync(&p, std::max<size_t>(bytes, 256), s)); }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p); }
    ~StreamBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

}  // namespace

void build_bvh_gpu(const float* pos, int num_tris, BvhBuild& out) {
    out = BvhBuild();
    if (num_tris <= 0) return;
    // Its own non-blocking stream: a build does not wait for, or hold up,
    // rendering on other streams.
    struct Stream {
        hipStream_t s = nullptr;
        Stream() { HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
        ~Stream() { (void)hipStreamDestroy(s); }
    } stream;
    hipStream_t st = stream.s;
    auto copy = [st](void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
        return e == hipSuccess ? hipStreamSynchronize(st) : e;
    };
    using Clock = std::chrono::steady_clock;
    auto ms_since = [](Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); };
    auto t_phase = Clock::now();
    const size_t n = (size_t)num_tris;
    const size_t cap = 2 * n;  // a binary tree over n leaves' ranges has < 2n nodes
    const size_t max_huge = n / kHugeNode + 1;
    const size_t max_tasks = n / kChunk + max_huge + 1;
    StreamBuf dpos(st), e0a(st), e0b(st), e0c(st), e1a(st), e1b(st), e1c(st), dorder(st), dscratch(st), nrange(st),
        nkids(st), nlo(st), nhi(st), dcount(st), dsmall(st), dmedium(st), dhuge(st), dacc(st), dtasks(st),
        dtask_first(st), dtask_cnt(st), dtask_off(st);
    dpos.alloc(n * 36);
    e0a.alloc(n * 16); e0b.alloc(n * 16); e0c.alloc(n * 8);
    e1a.alloc(n * 16); e1b.alloc(n * 16); e1c.alloc(n * 8);
    dorder.alloc(n * 4);
    dscratch.alloc(n * 4);
    nrange.alloc(cap * 16);
    nkids.alloc(cap * 8);
    nlo.alloc(cap * 16);
    nhi.alloc(cap * 16);
    dcount.alloc(256);
    dsmall.alloc(n * 4);  // a level has at most n nodes
    dmedium.alloc((n / kMediumNode + 1) * 4);
    dhuge.alloc(max_huge * 16);
    dacc.alloc(max_huge * sizeof(HugeAcc));
    dtasks.alloc(max_tasks * sizeof(HugeTask));
    dtask_first.alloc((max_huge + 1) * 4);
    dtask_cnt.alloc(max_tasks * 8);
    dtask_off.alloc(max_tasks * 8);
    HIP_OK(copy(dpos.p, pos, n * 36, hipMemcpyHostToDevice));
    Elems buf[2] = {{e0a.as<float4>(), e0b.as<float4>(), e0c.as<float2>()},
                    {e1a.as<float4>(), e1b.as<float4>(), e1c.as<float2>()}};
    Nodes nodes{nrange.as<int4>(), nkids.as<int2>(), nlo.as<float4>(), nhi.as<float4>()};
    k_bvh_init<<<(num_tris + 255) / 256, 256, 0, st>>>(dpos.as<float>(), num_tris, buf[0]);
    HIP_OK(hipGetLastError());
    int4 root = make_int4(0, num_tris, 0, 0);
    HIP_OK(copy(nrange.p, &root, sizeof root, hipMemcpyHostToDevice));
    // counters: nodes created, error bits, small / medium list lengths, fallback splits, huge list length
    int ctr[6] = {1, 0, 0, 0, 0, 0};
    HIP_OK(copy(dcount.p, ctr, sizeof ctr, hipMemcpyHostToDevice));
    int count = 1;
    HIP_OK(hipStreamSynchronize(st));
    out.gpu_phase_ms[0] = ms_since(t_phase);
    t_phase = Clock::now();
    int level_begin = 0, level_end = 1, cur = 0;
    std::vector<int> levels;  // first node of each level
    std::vector<int4> huge;
    std::vector<HugeTask> tasks;
    std::vector<int> task_first;
    while (level_begin < level_end) {
        levels.push_back(level_begin);
        const int nl = level_end - level_begin;
        k_bvh_classify<<<(nl + 255) / 256, 256, 0, st>>>(nodes.range, level_begin, level_end, dsmall.as<int32_t>(),
                                                  dmedium.as<int32_t>(), dhuge.as<int4>(), dcount.as<int>());
        HIP_OK(hipGetLastError());
        HIP_OK(copy(ctr, dcount.p, sizeof ctr, hipMemcpyDeviceToHost));
        const int nsmall = ctr[2], nmedium = ctr[3], nhuge = ctr[5];
        const Elems& src = buf[cur];
        const Elems& dst = buf[cur ^ 1];
        if (nhuge > 0) {
            huge.resize(nhuge);
            HIP_OK(copy(huge.data(), dhuge.p, (size_t)nhuge * 16, hipMemcpyDeviceToHost));
            tasks.clear();
            task_first.assign(nhuge + 1, 0);
            for (int j = 0; j < nhuge; ++j) {
                task_first[j] = (int)tasks.size();
                for (int c = huge[j].y; c < huge[j].z; c += kChunk)
                    tasks.push_back(HugeTask{j, c, std::min(c + kChunk, huge[j].z), huge[j].x});
            }
            task_first[nhuge] = (int)tasks.size();
            const int nt = (int)tasks.size();
            HIP_OK(copy(dtasks.p, tasks.data(), (size_t)nt * sizeof(HugeTask), hipMemcpyHostToDevice));
            HIP_OK(copy(dtask_first.p, task_first.data(), (size_t)(nhuge + 1) * 4, hipMemcpyHostToDevice));
            const HugeTask* tk = dtasks.as<HugeTask>();
            HugeAcc* acc = dacc.as<HugeAcc>();
            const int gs = (nhuge + 63) / 64;
            k_huge_init<<<gs, 64, 0, st>>>(acc, nhuge);
            k_huge_bounds<<<nt, kHugeThreads, 0, st>>>(src, tk, acc);
            k_huge_axis<<<gs, 64, 0, st>>>(src, dhuge.as<int4>(), nhuge, acc, nodes);
            k_huge_buckets<<<nt, kHugeThreads, 0, st>>>(src, tk, acc);
            k_huge_split<<<gs, 64, 0, st>>>(dhuge.as<int4>(), nhuge, acc, nodes, (int)cap, dcount.as<int>());
            k_huge_count<<<nt, kHugeThreads, 0, st>>>(src, tk, acc, dtask_cnt.as<int2>());
            k_huge_scan<<<gs, 64, 0, st>>>(dtask_first.as<int>(), nhuge, acc, dtask_cnt.as<int2>(), dtask_off.as<int2>(),
                                    dcount.as<int>());
            k_huge_move<true><<<nt, kHugeThreads, 0, st>>>(src, dst, dhuge.as<int4>(), tk, acc, dtask_off.as<int2>(),
                                                    dscratch.as<int32_t>());
            k_huge_move<false><<<nt, kHugeThreads, 0, st>>>(src, dst, dhuge.as<int4>(), tk, acc, dtask_off.as<int2>(),
                                                     dscratch.as<int32_t>());
            HIP_OK(hipGetLastError());
        }
        if (nmedium > 0)
            k_bvh_node<kHugeWaves><<<nmedium, kHugeThreads, 0, st>>>(src, dst, dorder.as<int32_t>(), dscratch.as<int32_t>(),
                                                              nodes, dmedium.as<int32_t>(), (int)cap,
                                                              dcount.as<int>());
        if (nsmall > 0)
            k_bvh_node<1><<<nsmall, kWave, 0, st>>>(src, dst, dorder.as<int32_t>(), dscratch.as<int32_t>(), nodes,
                                             dsmall.as<int32_t>(), (int)cap, dcount.as<int>());
        HIP_OK(hipGetLastError());
        HIP_OK(copy(ctr, dcount.p, sizeof ctr, hipMemcpyDeviceToHost));
        if (ctr[1] != 0 || (size_t)ctr[0] > cap) throw std::runtime_error("GPU BVH build: inconsistent partition");
        count = ctr[0];
        ctr[2] = ctr[3] = ctr[5] = 0;
        HIP_OK(copy(dcount.p, ctr, sizeof ctr, hipMemcpyHostToDevice));
        level_begin = level_end;
        level_end = count;
        cur ^= 1;
    }
    out.gpu_phase_ms[1] = ms_since(t_phase);
    t_phase = Clock::now();
    const int nn = count;
    const int depth = (int)levels.size() - 1;
    StreamBuf dsize(st), dicount(st), ddfs(st), dcomp(st), dlin(st), dpacked(st);
    dsize.alloc((size_t)nn * 4);
    dicount.alloc((size_t)nn * 4);
    ddfs.alloc((size_t)nn * 4);
    dcomp.alloc((size_t)nn * 4);
    levels.push_back(nn);
    for (int d = depth; d >= 0; --d) {
        int first = levels[d], last = levels[d + 1];
        k_bvh_sizes<<<(last - first + 255) / 256, 256, 0, st>>>(nodes.range, nodes.kids, first, last, dsize.as<int32_t>(),
                                                         dicount.as<int32_t>());
    }
    for (int d = 0; d <= depth; ++d) {
        int first = levels[d], last = levels[d + 1];
        k_bvh_index<<<(last - first + 255) / 256, 256, 0, st>>>(nodes.range, nodes.kids, first, last, dsize.as<int32_t>(),
                                                         dicount.as<int32_t>(), ddfs.as<int32_t>(),
                                                         dcomp.as<int32_t>());
    }
    int interior = 0;
    HIP_OK(copy(&interior, dicount.p, sizeof interior, hipMemcpyDeviceToHost));
    dlin.alloc((size_t)nn * sizeof(LinearNode));
    dpacked.alloc((size_t)std::max(interior, 1) * sizeof(PackedNode));
    k_bvh_emit<<<(nn + 255) / 256, 256, 0, st>>>(nodes, nn, ddfs.as<int32_t>(), dcomp.as<int32_t>(), dlin.as<LinearNode>(),
                                          dpacked.as<PackedNode>(), dcount.as<int>());
    HIP_OK(hipGetLastError());
    out.nodes.resize(nn);
    out.packed.resize(interior);
    out.order.resize(n);
    HIP_OK(copy(out.nodes.data(), dlin.p, (size_t)nn * sizeof(LinearNode), hipMemcpyDeviceToHost));
    if (interior) HIP_OK(copy(out.packed.data(), dpacked.p, (size_t)interior * sizeof(PackedNode),
                                   hipMemcpyDeviceToHost));
    HIP_OK(copy(out.order.data(), dorder.p, n * 4, hipMemcpyDeviceToHost));
    HIP_OK(copy(ctr, dcount.p, sizeof ctr, hipMemcpyDeviceToHost));
    out.gpu_phase_ms[2] = ms_since(t_phase);
    t_phase = Clock::now();
    out.fallback_splits = ctr[4];
    out.max_depth = depth;
    out.max_stack = depth + 1;
    // The root is a leaf only when the scene has at most four triangles.
    out.root_ref = interior ? 0u : leaf_ref(0u, (uint32_t)n);
    const LinearNode& root_node = out.nodes[0];
    out.bounds.lo = v3(root_node.lo[0], root_node.lo[1], root_node.lo[2]);
    out.bounds.hi = v3(root_node.hi[0], root_node.hi[1], root_node.hi[2]);
    out.gpu_phase_ms[3] = ms_since(t_phase);
}

}  // namespace rt
