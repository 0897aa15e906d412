// Wavefront path tracer for gfx950: kernels and the host frame driver.
//
// One render pass traces `batch` samples of every pixel at once (one path per
// (pixel, sample)); the bounce loop is
//     raygen -> [extend -> shade -> shadow] x max_bounces -> lost -> resolve
// with SoA float4 queues, persistent kernels that fetch 64-ray chunks from a
// device counter, wave-ballot compaction into the next queue, and device-side
// counts only (no host round trip inside a pass).
#include "render.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "engine.h"
#include "integrator.h"
#include "traverse.h"

namespace rt {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// Counter (kind, slice).
__device__ __forceinline__ int* ctr_at(int* counters, int kind, int slice) {
    return counters + ((size_t)kind * kSlices + slice) * kCounterStride;
}
__device__ __forceinline__ int load_relaxed(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lanemask_lt() {
    int lane = threadIdx.x & (kWave - 1);
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
// Reserve slots for the lanes with `pred` set; returns this lane's slot.
__device__ __forceinline__ int wave_append(bool pred, int* ctr) {
    uint64_t m = __ballot(pred);
    int n = __popcll(m);
    int base = 0;
    if ((threadIdx.x & (kWave - 1)) == 0 && n > 0) base = atomicAdd(ctr, n);
    base = __shfl(base, 0);
    return base + __popcll(m & lanemask_lt());
}
__device__ __forceinline__ void wave_add_u64(unsigned long long* p, uint32_t v) {
    // Wave reduction then one atomic per wave.
    uint64_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor((unsigned long long)s, off);
    if ((threadIdx.x & (kWave - 1)) == 0 && s) atomicAdd(p, (unsigned long long)s);
}

// Persistent-wave work distribution over the slices of one list: a wave starts
// on its own slice and moves on (stealing) when a slice is drained.  Each call
// to next() yields a 64-item chunk (slice, first item) or false when every
// slice is drained.  All lanes of the wave see the same values.
//
// When its slice runs dry the wave looks at every slice at once -- lane k
// reads slice k's count and fetch counter, one round trip instead of one per
// slice -- and exits when none is live.  Otherwise it steals from a live
// slice of its own XCD if there is one: workgroups are dispatched to the 8
// XCDs round-robin, so with 4 waves per block the waves that start on slice
// s run on XCD (s >> 2) & 7, and a thief from the same XCD finds the slice's
// BVH nodes in its own L2.  Among the candidates, a per-wave rotation picks
// the first live one, which spreads thieves over the live slices.  Counts
// are fixed for the launch (written by the previous stage) and fetch counters
// only grow, so a slice seen drained stays drained and every wave's loop ends.
// `limit` caps every slice's count (a list whose appends may overrun its
// per-slice room holds only the first `limit` entries).
static_assert(kSlices == kWave, "the drained-slice scan maps one slice per lane");
static_assert(kBlock / kWave == 4 && kSlices == 64, "slice-to-XCD map assumes 4 waves per block, 64 slices");
struct SliceCursor {
    int* counters;
    int cnt_kind, fetch_kind;
    int slice, count, rot;
    uint64_t own;  // slices whose first waves run on this wave's XCD
    int limit;
    __device__ void init(int* c, int ck, int fk, int lim = 0x7fffffff) {
        counters = c;
        cnt_kind = ck;
        fetch_kind = fk;
        limit = lim;
        int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
        slice = wave % kSlices;
        rot = (wave / kSlices) % kSlices;
        const int xcd = (int)(blockIdx.x & 7u);
        own = (0xFull << (4 * xcd)) | (0xFull << (32 + 4 * xcd));
        count = min(*ctr_at(counters, cnt_kind, slice), limit);
    }
    __device__ bool next(int& base) {
        const int lane = (int)(threadIdx.x & (kWave - 1));
        while (true) {
            int* f = ctr_at(counters, fetch_kind, slice);
            int b = count;
            // Lane 0 claims the chunk and broadcasts, so the wave stays uniform.
            if (lane == 0 && load_relaxed(f) < count) b = atomicAdd(f, kWave);
            b = __shfl(b, 0);
            if (b < count) {
                base = b;
                return true;
            }
            // Drained: lane k checks slice k.
            const int c = min(*ctr_at(counters, cnt_kind, lane), limit);
            const bool live = load_relaxed(ctr_at(counters, fetch_kind, lane)) < c;
            const uint64_t m = __ballot(live);
            if (m == 0) return false;
            const uint64_t cand = (m & own) ? (m & own) : m;
            // First candidate at or after slice `rot` (cyclically).
            const uint64_t hi = cand & (~0ull << rot);
            const int k = __ffsll((unsigned long long)(hi ? hi : cand)) - 1;
            slice = k;
            count = __shfl(c, k);
        }
    }
};

// ------------------------------------------------------------------ jobs
// POOLED kernels serve passes of many jobs: the job index is read from the
// path's flags word / shadow tag and selects the scene and frame constants.
// Single-job passes instantiate POOLED = false, where the job is 0 and every
// job field is a uniform (scalar) load.
template <bool POOLED>
__device__ __forceinline__ const JobDev& pass_job(const JobDev* __restrict__ jobs, int job) {
    return jobs[POOLED ? job : 0];
}
constexpr uint32_t kFlagMask = (1u << kJobShift) - 1u;

// Job owning pass path p (paths of a job are contiguous, bases ascending).
__device__ __forceinline__ int job_of_path(const JobDev* __restrict__ jobs, int n_jobs, int p) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].path_base <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int job_of_pixel(const JobDev* __restrict__ jobs, int n_jobs, int g) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].pix_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ raygen
// Jittered pinhole ray per (pixel, sample) (reference: Src/Core/camera.cu:27-35,
// Src/Core/renderer.cuh:96-115).  Local path l of a job covers pixel l % wh of
// sample sample0 + l / wh; row 0 of the image is the bottom row.  Pass path p
// sits at queue index p, in slice p / region.
template <bool POOLED>
__global__ __launch_bounds__(kBlock) void k_raygen(const PassHeader* __restrict__ hdr, const JobDev* __restrict__ jobs,
                                                    float4* qo, float4* qd, float4* qw, float4* slot_d, float4* slot_i,
                                                    int* counters, int qslot) {
    const int n_paths = hdr->n_paths;
    const int n_jobs = POOLED ? hdr->n_jobs : 1;
    const int region = hdr->region;
    int stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        int n = n_paths - sl * region;
        *ctr_at(counters, qslot, sl) = n < 0 ? 0 : (n > region ? region : n);
        *ctr_at(counters, kFetchExtend, sl) = 0;
        *ctr_at(counters, kCntCont, sl) = 0;
        *ctr_at(counters, kFetchCont, sl) = 0;
        for (int c = 0; c < kNumClasses; ++c) {
            *ctr_at(counters, kCntClass0 + c, sl) = 0;
            *ctr_at(counters, kFetchClass0 + c, sl) = 0;
        }
    }
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_paths; p += stride) {
        const int job = POOLED ? job_of_path(jobs, n_jobs, p) : 0;
        const JobDev& J = pass_job<POOLED>(jobs, job);
        const int wh = J.fp.wh, width = J.fp.width;
        const int local = p - J.path_base;
        uint32_t pixel = (uint32_t)(local % wh);
        uint32_t sample = J.fp.sample0 + (uint32_t)(local / wh);
        int i = (int)(pixel % (uint32_t)width), j = (int)(pixel / (uint32_t)width);
        Rng rng;
        rng.init(J.fp.seed, pixel, sample, 0u, kStreamCamera);
        float x0 = rng.next();
        float x1 = rng.next();
        const vec3 cam = J.fp.cam_pos;
        vec3 pp = (J.fp.left_down + ((float)i + x0) * J.fp.du) + ((float)j + x1) * J.fp.dv;
        vec3 d = normalize(pp - cam);
        qo[p] = make_float4(cam.x, cam.y, cam.z, 0.f);
        qd[p] = make_float4(d.x, d.y, d.z, i_as_f((int)(kFlagFirst | ((uint32_t)job << kJobShift))));
        qw[p] = make_float4(1.f, 1.f, 1.f, i_as_f(p));
        slot_d[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        slot_i[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ------------------------------------------------------------------ ray order
// Optional reordering of each slice's extension rays before extend (render
// setting ray_sort).  Shade writes a 9-bit key beside every ray it appends:
// the direction's octant and the origin's cell in a 4x4x4 grid over the job's
// scene bounds (Morton order).  The keys are counting-sorted per slice into a
// permutation that extend reads its rays through, so the rays of a wave tend
// to share their octant and region and take similar paths through the tree.
// Extend writes each hit, and a copy of the ray, at the sorted position, so
// shade reads them in order; only the path weight stays in its queue slot.
// Every path's arithmetic is the same in any order: results do not change.
constexpr int kSortBins = 512;

__device__ __forceinline__ uint32_t ray_sort_key(vec3 o, vec3 d, const DevScene& sc) {
    const uint32_t oct = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
    auto cell = [](float x, float lo, float hi) {
        const int c = (int)((x - lo) * (4.f / fmaxf(hi - lo, 1e-30f)));
        return (uint32_t)min(max(c, 0), 3);
    };
    const uint32_t cx = cell(o.x, sc.bounds_lo.x, sc.bounds_hi.x);
    const uint32_t cy = cell(o.y, sc.bounds_lo.y, sc.bounds_hi.y);
    const uint32_t cz = cell(o.z, sc.bounds_lo.z, sc.bounds_hi.z);
    const uint32_t m = (cx & 1u) | ((cy & 1u) << 1) | ((cz & 1u) << 2) | ((cx & 2u) << 2) | ((cy & 2u) << 3) |
                       ((cz & 2u) << 4);
    return (oct << 6) | m;
}

constexpr int kSortItems = 8;  // rays per thread per scatter round

__global__ __launch_bounds__(kBlock) void k_sort_count(const uint16_t* __restrict__ keys,
                                                        const PassHeader* __restrict__ ph, int* counters, int qslot,
                                                        int* hist) {
    __shared__ int h[kSortBins];
    const int sl = blockIdx.y, region = ph->region;
    const int n = *ctr_at(counters, qslot, sl);
    for (int t = threadIdx.x; t < kSortBins; t += kBlock) h[t] = 0;
    __syncthreads();
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        atomicAdd(&h[keys[sl * region + j]], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < kSortBins; t += kBlock)
        if (h[t]) atomicAdd(&hist[sl * kSortBins + t], h[t]);
}

// Exclusive scan of each slice's histogram into the scatter cursors; the
// histogram is cleared for the next bounce.  One block of kSortBins threads per slice.
__global__ __launch_bounds__(kSortBins) void k_sort_scan(int* hist, int* cur) {
    __shared__ int sh[kSortBins];
    const int sl = blockIdx.x, t = threadIdx.x;
    const int v = hist[sl * kSortBins + t];
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kSortBins; off <<= 1) {
        const int a = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    cur[sl * kSortBins + t] = sh[t] - v;
    hist[sl * kSortBins + t] = 0;
}

// Positions: a block ranks kSortItems * kBlock keys in LDS, reserves each
// bin's run with one global add, and writes the queue positions into the
// slice's permutation.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const uint16_t* __restrict__ keys,
                                                          const PassHeader* __restrict__ ph, int* counters, int qslot,
                                                          int* cur, int* __restrict__ perm) {
    __shared__ int h[kSortBins], base[kSortBins];
    const int sl = blockIdx.y, region = ph->region;
    const int n = *ctr_at(counters, qslot, sl);
    for (int t = threadIdx.x; t < kSortBins; t += kBlock) h[t] = 0;
    __syncthreads();
    constexpr int CHUNK = kSortItems * kBlock;
    for (int j0 = blockIdx.x * CHUNK; j0 < n; j0 += gridDim.x * CHUNK) {  // uniform per block
        uint32_t key[kSortItems];
        int r[kSortItems];
#pragma unroll
        for (int k = 0; k < kSortItems; ++k) {
            const int j = j0 + k * kBlock + (int)threadIdx.x;
            key[k] = 0;
            r[k] = -1;
            if (j < n) {
                key[k] = keys[sl * region + j];
                r[k] = atomicAdd(&h[key[k]], 1);
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kSortBins; t += kBlock)
            if (h[t]) base[t] = atomicAdd(&cur[sl * kSortBins + t], h[t]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortItems; ++k)
            if (r[k] >= 0) perm[sl * region + base[key[k]] + r[k]] = j0 + k * kBlock + (int)threadIdx.x;
        __syncthreads();
        for (int t = threadIdx.x; t < kSortBins; t += kBlock) h[t] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ extend
// Closest hit for every path of the current queue, then the path's queue
// index is appended to its slice's list for the shading class of what it hit.
// A miss ends its path here instead: the environment's radiance times the
// path weight (no MIS, reference: Src/Core/renderer.cuh:142-159) is added to
// the path's sum at once, so the shade stage never sees misses, and their
// texel reads overlap the traversal of the other waves.
// The traversal stacks live in dynamic LDS: one column of `max stack depth`
// entries per lane, sized per pass to the deepest tree among its jobs, so a
// shallow tree leaves room for more resident waves.
extern __shared__ uint32_t stk_dyn[];

// Extension rays whose search is suspended (the stragglers of a wave, see
// trace_closest_run) are stored per slice for k_extend_cont: record k of
// slice s sits at s * room + k; its hit so far and (launch position, queue
// slot, next item, stack depth) are one float4 and one int4, its stack entry
// e is word e * total + s * room + k.  Appends beyond a slice's room are not
// stored: those lanes finish their search in k_extend.
struct ContBufs {
    float4* hit;
    int4* state;
    uint32_t* stack;
    int room;      // records per slice
    size_t total;  // room * kSlices
};
// A wave's searching lanes suspend once fewer than this many are left; passes
// whose trees need fewer than kSuspendMinStack stack entries (a few thousand
// triangles at most) do not suspend, their searches are short.
constexpr int kSuspendLanes = 12;
constexpr int kSuspendMinStack = 12;
// Shadow rays end at the light point, so their searches are shorter: they
// suspend only in deeper trees (a 20 k-triangle scene of depth 18 loses 12 %
// of its shadow time to the continuation launch, a 2 M-triangle terrain of
// depth 26 gains 8 %).
constexpr int kShadowSuspendMinStack = 24;
// Each suspending stage adds a launch per bounce, which a small pass (a
// real-time frame of 0.9 M paths: +0.5 ms per frame for each stage) does not
// earn back; passes of fewer paths than this do not suspend.
constexpr size_t kSuspendMinPaths = (size_t)1 << 24;
static_assert(kContStack == (size_t)kSuspendMaxStack, "stored stack depth");

// What extend does with a finished search: the hit record, the ray copied to
// its sorted position (after a ray sort), the shading class; a miss adds the
// environment's emission to the path's sums.  cls = -1 when nothing is appended.
template <bool POOLED>
__device__ __forceinline__ void extend_finish(const DevScene& sc, const Hit& h, int p, float4 o, float4 d, int i,
                                              const float4* __restrict__ qw, float4* __restrict__ hits,
                                              float4* slot_d, float4* slot_i, const int* __restrict__ perm,
                                              float4* __restrict__ sorted_o, float4* __restrict__ sorted_d,
                                              int& cls, bool& miss) {
    miss = h.prim < 0;
    if (!miss) {
        hits[p] = make_float4(h.t, h.b1, h.b2, i_as_f(h.prim));
        if (perm) {
            sorted_o[p] = o;
            sorted_d[p] = d;
        }
        cls = shade_class_of(sc.tri_attr, h.prim);
    } else if (sc.env_light >= 0) {
        const float4 w = qw[i];
        const vec3 e = cmul(v3(w.x, w.y, w.z), env_emission(sc, v3(d.x, d.y, d.z)));
        float4* sum = (f_as_i(d.w) & kFlagFirst) ? slot_d : slot_i;
        const int pid = f_as_i(w.w);
        float4 a = sum[pid];
        a.x += e.x;
        a.y += e.y;
        a.z += e.z;
        sum[pid] = a;
    }
}

// Class-list appends and counters of one wave's finished rays.
template <bool COUNT>
__device__ __forceinline__ void extend_append(int cls, bool miss, int p, int slice, int region, int* lists, int cap,
                                              int* counters, unsigned long long* stats, const TravCount& c) {
#pragma unroll
    for (int k = 0; k < kNumClasses; ++k) {
        if (__ballot(cls == k) == 0ull) continue;
        int pos = wave_append(cls == k, ctr_at(counters, kCntClass0 + k, slice));
        if (cls == k) lists[(size_t)k * cap + (size_t)slice * region + pos] = p;
    }
    if (COUNT) {
        wave_add_u64(&stats[kStatExtNodes], c.nodes);
        wave_add_u64(&stats[kStatExtTris], c.tris);
        const uint64_t mm = __ballot(miss);
        if ((threadIdx.x & (kWave - 1)) == 0 && mm) atomicAdd(&stats[kStatMisses], (unsigned long long)__popcll(mm));
#pragma unroll
        for (int k = 0; k < kNumClasses; ++k) {
            const uint64_t m = __ballot(cls == k);
            if ((threadIdx.x & (kWave - 1)) == 0 && m)
                atomicAdd(&stats[kStatClassPaths + k], (unsigned long long)__popcll(m));
        }
    }
}

template <bool COUNT, bool POOLED, int SUSP>
__global__ __launch_bounds__(kBlock) void k_extend(const JobDev* __restrict__ jobs, const float4* __restrict__ qo,
                                                    const float4* __restrict__ qd, const float4* __restrict__ qw,
                                                    float4* __restrict__ hits, int* __restrict__ lists, int cap,
                                                    float4* slot_d, float4* slot_i,
                                                    const PassHeader* __restrict__ ph, int* counters, int qslot,
                                                    unsigned long long* stats, const int* __restrict__ perm,
                                                    float4* __restrict__ sorted_o, float4* __restrict__ sorted_d,
                                                    ContBufs cont) {
    const int region = ph->region;
    uint32_t* stk = stk_dyn;
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        *ctr_at(counters, qslot ^ 1, sl) = 0;
        *ctr_at(counters, kCntShadow, sl) = 0;
        *ctr_at(counters, kFetchShadow, sl) = 0;
        *ctr_at(counters, kCntShadowCont, sl) = 0;
        *ctr_at(counters, kFetchShadowCont, sl) = 0;
        int n = *ctr_at(counters, qslot, sl);
        if (n > 0) atomicAdd(&stats[kStatExtRays], (unsigned long long)n);
    }
    uint32_t* col = stk + threadIdx.x;
    SliceCursor cur;
    cur.init(counters, qslot, kFetchExtend);
    int base;
    while (cur.next(base)) {
        int j = base + (int)(threadIdx.x & (kWave - 1));
        // p: the position in this launch's order, where the hit, the class list
        // entry and (after a ray sort) a copy of the ray go; i: the ray's queue slot.
        const int p = cur.slice * region + j;
        int i = p;
        TravCount c{0, 0};
        int cls = -1;
        bool miss = false, susp = false;
        const bool live = j < cur.count;
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
        if (live) {
            if (perm) i = cur.slice * region + perm[p];
            o = qo[i];
            d = qd[i];
        }
        const DevScene& sc = pass_job<POOLED>(jobs, live ? (int)((uint32_t)f_as_i(d.w) >> kJobShift) : 0).sc;
        const RayInv r = make_ray(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
        TravState st = trav_start(sc.root_ref, 1e8f);
        if (live) susp = trace_closest_run_wave<COUNT, true, false, SUSP>(sc, r, 0.f, col, kBlock, c, st);
        if (SUSP > 0 && __ballot(susp) != 0ull) {
            const int pos = wave_append(susp, ctr_at(counters, kCntCont, cur.slice));
            if (susp) {
                if (pos < cont.room) {
                    const size_t k = (size_t)cur.slice * cont.room + pos;
                    cont.hit[k] = make_float4(st.h.t, st.h.b1, st.h.b2, i_as_f(st.h.prim));
                    cont.state[k] = make_int4(p, i, (int)st.ref, st.sp);
                    for (int e = 0; e < st.sp; ++e) cont.stack[(size_t)e * cont.total + k] = col[e * kBlock];
                } else {
                    trace_closest_run<COUNT>(sc.nodes, sc.tri_isect, r, 0.f, col, kBlock, c, st);
                    susp = false;
                }
            }
        }
        if (live && !susp) extend_finish<POOLED>(sc, st.h, p, o, d, i, qw, hits, slot_d, slot_i, perm, sorted_o,
                                                 sorted_d, cls, miss);
        extend_append<COUNT>(cls, miss, p, cur.slice, region, lists, cap, counters, stats, c);
    }
}

// The suspended searches of k_extend, gathered 64 to a wave: each lane loads
// a record, puts its stack back into its LDS column and resumes the search
// where it stopped (the same tests in the same order, so the same hit), then
// finishes the ray as k_extend does, into the same slice's lists.
template <bool COUNT, bool POOLED>
__global__ __launch_bounds__(kBlock) void k_extend_cont(const JobDev* __restrict__ jobs,
                                                         const float4* __restrict__ qo, const float4* __restrict__ qd,
                                                         const float4* __restrict__ qw, float4* __restrict__ hits,
                                                         int* __restrict__ lists, int cap, float4* slot_d,
                                                         float4* slot_i, const PassHeader* __restrict__ ph,
                                                         int* counters, unsigned long long* stats,
                                                         const int* __restrict__ perm, float4* __restrict__ sorted_o,
                                                         float4* __restrict__ sorted_d, ContBufs cont) {
    const int region = ph->region;
    uint32_t* col = stk_dyn + threadIdx.x;
    SliceCursor cur;
    cur.init(counters, kCntCont, kFetchCont, cont.room);
    int base;
    while (cur.next(base)) {
        const int j = base + (int)(threadIdx.x & (kWave - 1));
        TravCount c{0, 0};
        int cls = -1, p = 0;
        bool miss = false;
        if (j < cur.count) {
            const size_t k = (size_t)cur.slice * cont.room + j;
            const int4 s = cont.state[k];
            const float4 hh = cont.hit[k];
            p = s.x;
            const int i = s.y;
            for (int e = 0; e < s.w; ++e) col[e * kBlock] = cont.stack[(size_t)e * cont.total + k];
            const float4 o = qo[i], d = qd[i];
            const DevScene& sc = pass_job<POOLED>(jobs, (int)((uint32_t)f_as_i(d.w) >> kJobShift)).sc;
            const RayInv r = make_ray(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
            TravState st;
            st.h.t = hh.x;
            st.h.b1 = hh.y;
            st.h.b2 = hh.z;
            st.h.prim = f_as_i(hh.w);
            st.ref = (uint32_t)s.z;
            st.sp = s.w;
            trace_closest_run_wave<COUNT>(sc, r, 0.f, col, kBlock, c, st);
            extend_finish<POOLED>(sc, st.h, p, o, d, i, qw, hits, slot_d, slot_i, perm, sorted_o, sorted_d, cls,
                                  miss);
        }
        extend_append<COUNT>(cls, miss, p, cur.slice, region, lists, cap, counters, stats, c);
    }
}

// Shadow records are stored in blocks of 64: field f (0..3) of record r sits
// at shadow[(r / 64) * 256 + f * 64 + r % 64], so a wave's 64 consecutive
// records write and read each field as one contiguous 1 KiB run (two runs when
// they straddle a block) instead of 16 bytes at a 64-byte stride.  The buffer
// holds the engine's capacity rounded up to whole blocks.
template <typename T>
__device__ __forceinline__ T* shadow_field(T* base, size_t r, int f) {
    return base + ((r >> 6) << 8) + ((size_t)f << 6) + (r & 63);
}

// ------------------------------------------------------------------ shade
// One launch per bounce shades every class present in the pass: its waves
// drain the class lists (binned by extend) in turn, the subsurface class
// first -- few paths, each tracing two probe rays, a long latency that the
// other classes' work then overlaps instead of running as a launch of its own
// with the GPU nearly idle -- and the diffuse class, the bulk, last.  Each
// class's code is specialised (shade_class<CLS>), so a wave never mixes
// materials.  Survivors and shadow records are appended to the slice they
// came from.  At a job's last bounce its survivors are dropped together with
// everything they gathered (reference: Src/Core/renderer.cuh:1166-1194 never
// accumulates a path still alive after maxBounces), unless the job flushes
// survivors.
template <int CLS, bool POOLED>
__device__ __forceinline__ void shade_class(const JobDev* __restrict__ jobs, int bounce,
                                            const float4* __restrict__ qo, const float4* __restrict__ qd,
                                            const float4* __restrict__ qw, const float4* __restrict__ hits,
                                            const int* __restrict__ list, float4* __restrict__ no,
                                            float4* __restrict__ nd, float4* __restrict__ nw,
                                            float4* __restrict__ shadow, float4* slot_d, float4* slot_i,
                                            int region, int* counters, int qslot, uint16_t* __restrict__ nkey,
                                            const int* __restrict__ perm, const float4* __restrict__ sorted_o,
                                            const float4* __restrict__ sorted_d, uint32_t* col,
                                            unsigned long long* stats) {
    SliceCursor cur;
    cur.init(counters, kCntClass0 + CLS, kFetchClass0 + CLS);
    int base;
    while (cur.next(base)) {
        int j = base + (int)(threadIdx.x & (kWave - 1));
        ShadeOut so;
        so.alive = false;
        so.shadow = false;
        so.add_direct = false;
        so.sss_rays = 0;
        PathState ps;
        int pid = 0;
        uint32_t jobbits = 0;
        uint32_t key = 0;
        if (j < cur.count) {
            // List entries are extend's positions: after a ray sort the ray's
            // origin and direction were copied there and its weight stays in
            // its queue slot; otherwise position and slot are the same.
            const int i = list[(size_t)cur.slice * region + j];
            float4 o, d, w;
            if (perm) {
                o = sorted_o[i];
                d = sorted_d[i];
                w = qw[cur.slice * region + perm[i]];
            } else {
                o = qo[i];
                d = qd[i];
                w = qw[i];
            }
            const uint32_t flags = (uint32_t)f_as_i(d.w);
            jobbits = flags & ~kFlagMask;
            const JobDev& J = pass_job<POOLED>(jobs, (int)(flags >> kJobShift));
            FrameParams fp = J.fp;
            fp.bounce = bounce;
            ps.o = v3(o.x, o.y, o.z);
            ps.pdf = o.w;
            ps.d = v3(d.x, d.y, d.z);
            ps.flags = flags & kFlagMask;
            ps.w = v3(w.x, w.y, w.z);
            pid = f_as_i(w.w);
            const int local = pid - J.path_base;
            uint32_t pixel = (uint32_t)(local % fp.wh);
            uint32_t sample = fp.sample0 + (uint32_t)(local / fp.wh);
            so = shade_path<CLS>(J.sc, fp, ps, hits[i], pixel, sample, col, kBlock);
            if (so.has_emit) {
                float4* s = so.add_direct ? slot_d : slot_i;
                float4 a = s[pid];
                a.x += so.emit.x;
                a.y += so.emit.y;
                a.z += so.emit.z;
                s[pid] = a;
            }
            if (CLS != kClsTerminal && so.alive && bounce + 1 >= fp.max_bounces) {
                so.alive = false;
                if (!J.flush_survivors) {
                    slot_d[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
                    slot_i[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
                    so.shadow = false;
                }
            }
            if (CLS != kClsTerminal && nkey && so.alive) key = ray_sort_key(ps.o, ps.d, J.sc);
        }
        if (CLS == kClsSubsurface) {
            // Rays traced by the subsurface events (probe rays, walk segments): one add per wave.
            int n = so.sss_rays;
            for (int off = kWave / 2; off > 0; off >>= 1) n += __shfl_xor(n, off);
            if ((threadIdx.x & (kWave - 1)) == 0 && n) atomicAdd(&stats[kStatProbeRays], (unsigned long long)n);
        }
        if (CLS != kClsTerminal) {
            int qpos = wave_append(so.alive, ctr_at(counters, qslot ^ 1, cur.slice));
            if (so.alive) {
                size_t o = (size_t)cur.slice * region + qpos;
                no[o] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.pdf);
                nd[o] = make_float4(ps.d.x, ps.d.y, ps.d.z, i_as_f((int)((ps.flags & kFlagMask) | jobbits)));
                nw[o] = make_float4(ps.w.x, ps.w.y, ps.w.z, i_as_f(pid));
                if (nkey) nkey[o] = (uint16_t)key;
            }
        }
        if (CLS == kClsDiffuse || CLS == kClsPbr || CLS == kClsSubsurface) {
            int spos = wave_append(so.shadow, ctr_at(counters, kCntShadow, cur.slice));
            if (so.shadow) {
                // tag = job << 1 | (sum: 0 direct, 1 indirect)
                int tag = (int)((jobbits >> kJobShift) << 1) | (so.add_direct ? 0 : 1);
                const size_t r = (size_t)cur.slice * region + spos;
                *shadow_field(shadow, r, 0) = make_float4(so.rec.lp.x, so.rec.lp.y, so.rec.lp.z, so.rec.sdist);
                *shadow_field(shadow, r, 1) = make_float4(so.rec.dir.x, so.rec.dir.y, so.rec.dir.z, i_as_f(pid));
                *shadow_field(shadow, r, 2) = make_float4(so.rec.hp.x, so.rec.hp.y, so.rec.hp.z, i_as_f(tag));
                *shadow_field(shadow, r, 3) =
                    make_float4(so.rec.contrib.x, so.rec.contrib.y, so.rec.contrib.z, 0.f);
            }
        }
    }
}

// classes: bit c set when class c occurs in the pass.  The stack columns in
// dynamic LDS are sized only when the subsurface class does (its probe rays).
template <bool POOLED>
__global__ __launch_bounds__(kBlock) void k_shade(const JobDev* __restrict__ jobs, int bounce, int classes,
                                                   const float4* __restrict__ qo, const float4* __restrict__ qd,
                                                   const float4* __restrict__ qw, const float4* __restrict__ hits,
                                                   const int* __restrict__ lists, size_t cap,
                                                   float4* __restrict__ no, float4* __restrict__ nd,
                                                   float4* __restrict__ nw, float4* __restrict__ shadow,
                                                   float4* slot_d, float4* slot_i,
                                                   const PassHeader* __restrict__ ph, int* counters, int qslot,
                                                   uint16_t* __restrict__ nkey, const int* __restrict__ perm,
                                                   const float4* __restrict__ sorted_o,
                                                   const float4* __restrict__ sorted_d,
                                                   unsigned long long* stats) {
    const int region = ph->region;
    uint32_t* col = stk_dyn + threadIdx.x;
    auto run = [&](auto cls) {
        constexpr int CLS = decltype(cls)::value;
        if (classes & (1 << CLS))
            shade_class<CLS, POOLED>(jobs, bounce, qo, qd, qw, hits, lists + (size_t)CLS * cap, no, nd, nw, shadow,
                                     slot_d, slot_i, region, counters, qslot, nkey, perm, sorted_o, sorted_d, col,
                                     stats);
    };
    run(std::integral_constant<int, kClsSubsurface>());
    run(std::integral_constant<int, kClsDielectric>());
    run(std::integral_constant<int, kClsPbr>());
    run(std::integral_constant<int, kClsTerminal>());
    run(std::integral_constant<int, kClsDiffuse>());
}

// ------------------------------------------------------------------ shadow
// MODE (rt_render_params.shadow_any_hit):
// 0: the reference's test (Src/Core/light.cu:5-27): the closest hit from the
//    light point towards the shading point with tMin 1e-6, then its 0.1
//    distance tolerances.
// 1: the same ray, stopped at the first hit clearly in front of the shading
//    point, which the closest-hit test would call occluded whenever the
//    closest hit lies on the ray where its t says (see trace_closest's EARLY).
// 2: the segment test: from the shading point towards the light point, any
//    triangle the reference's ray would meet from the front, between 0.1 past
//    the shading point and 0.01 short of the light point (the reference's
//    tolerances as distances along the segment), occludes.  Hits beyond the
//    shading point and at the shading point's distance from the light, which
//    the reference's test also calls occluded, are not looked for: an
//    approximation of mode 0 whose difference the tests bound.
constexpr int kShadowClosest = 0, kShadowEarly = 1, kShadowSegment = 2;
// The reference's visibility decision from a finished closest-hit search of a
// shadow record's ray; a visible light point adds the record's contribution
// to its path's direct or indirect sum.
__device__ __forceinline__ bool shadow_finish(const DevScene& sc, const float4& r0, const float4& r1, const float4& r2,
                                              const float4& r3, const Hit& h, float4* slot_d, float4* slot_i) {
    ShadowRec rec;
    rec.lp = v3(r0.x, r0.y, r0.z);
    rec.sdist = r0.w;
    rec.dir = v3(r1.x, r1.y, r1.z);
    rec.hp = v3(r2.x, r2.y, r2.z);
    const bool hit = h.prim >= 0;
    vec3 tp = v3(0.f, 0.f, 0.f);
    if (hit) tp = position_at(sc.tri_attr, h.prim, h.b1, h.b2);
    if (h.prim == kEarlyExit || !shadow_visible(rec, hit, tp)) return false;
    float4* s = (f_as_i(r2.w) & 1) == 0 ? slot_d : slot_i;
    const int pid = f_as_i(r1.w);
    float4 a = s[pid];
    a.x += r3.x;
    a.y += r3.y;
    a.z += r3.z;
    s[pid] = a;
    return true;
}

// SUSP > 0 (closest-hit mode only): stragglers are suspended as in k_extend
// (their record index instead of the ray) and k_shadow_cont finishes them.
template <bool COUNT, bool POOLED, int MODE = kShadowClosest, int SUSP = 0>
__global__ __launch_bounds__(kBlock) void k_shadow(const JobDev* __restrict__ jobs, const float4* __restrict__ shadow,
                                                    float4* slot_d, float4* slot_i, const PassHeader* __restrict__ ph,
                                                    int* counters, unsigned long long* stats, ContBufs cont) {
    static_assert(SUSP == 0 || MODE == kShadowClosest, "only the closest-hit test suspends");
    const int region = ph->region;
    uint32_t* stk = stk_dyn;
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        int n = *ctr_at(counters, kCntShadow, sl);
        if (n > 0) atomicAdd(&stats[kStatShadowRays], (unsigned long long)n);
        // The queue and the class lists of this bounce are consumed: reset their
        // fetch counters and counts for the next extend.
        *ctr_at(counters, kFetchExtend, sl) = 0;
        *ctr_at(counters, kCntCont, sl) = 0;
        *ctr_at(counters, kFetchCont, sl) = 0;
        for (int c = 0; c < kNumClasses; ++c) {
            *ctr_at(counters, kCntClass0 + c, sl) = 0;
            *ctr_at(counters, kFetchClass0 + c, sl) = 0;
        }
    }
    uint32_t* col = stk + threadIdx.x;
    SliceCursor cur;
    cur.init(counters, kCntShadow, kFetchShadow);
    int base;
    while (cur.next(base)) {
        int j = base + (int)(threadIdx.x & (kWave - 1));
        TravCount c{0, 0};
        bool lit = false;
        if (SUSP > 0) {
            const bool live = j < cur.count;
            const size_t r = (size_t)cur.slice * region + j;
            float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0, r3 = r0;
            if (live) {
                r0 = *shadow_field(shadow, r, 0);
                r1 = *shadow_field(shadow, r, 1);
                r2 = *shadow_field(shadow, r, 2);
                r3 = *shadow_field(shadow, r, 3);
            }
            const DevScene& sc = pass_job<POOLED>(jobs, live ? f_as_i(r2.w) >> 1 : 0).sc;
            const vec3 dir = v3(r1.x, r1.y, r1.z);
            const RayInv ray = make_ray(v3(r0.x, r0.y, r0.z) + 0.01f * dir, dir);
            TravState st = trav_start(sc.root_ref, 1e8f);
            bool susp = false;
            if (live) susp = trace_closest_run_wave<COUNT, true, false, SUSP>(sc, ray, 1e-6f, col, kBlock, c, st);
            if (__ballot(susp) != 0ull) {
                const int pos = wave_append(susp, ctr_at(counters, kCntShadowCont, cur.slice));
                if (susp) {
                    if (pos < cont.room) {
                        const size_t k = (size_t)cur.slice * cont.room + pos;
                        cont.hit[k] = make_float4(st.h.t, st.h.b1, st.h.b2, i_as_f(st.h.prim));
                        cont.state[k] = make_int4((int)r, 0, (int)st.ref, st.sp);
                        for (int e = 0; e < st.sp; ++e) cont.stack[(size_t)e * cont.total + k] = col[e * kBlock];
                    } else {
                        trace_closest_run<COUNT>(sc.nodes, sc.tri_isect, ray, 1e-6f, col, kBlock, c, st);
                        susp = false;
                    }
                }
            }
            if (live && !susp) lit = shadow_finish(sc, r0, r1, r2, r3, st.h, slot_d, slot_i);
        } else if (j < cur.count) {
            const size_t r = (size_t)cur.slice * region + j;
            const float4 r0 = *shadow_field(shadow, r, 0), r1 = *shadow_field(shadow, r, 1);
            const float4 r2 = *shadow_field(shadow, r, 2), r3 = *shadow_field(shadow, r, 3);
            const int tag = f_as_i(r2.w);
            const DevScene& sc = pass_job<POOLED>(jobs, tag >> 1).sc;
            ShadowRec rec;
            rec.lp = v3(r0.x, r0.y, r0.z);
            rec.sdist = r0.w;
            rec.dir = v3(r1.x, r1.y, r1.z);
            rec.hp = v3(r2.x, r2.y, r2.z);
            vec3 org = rec.lp + 0.01f * rec.dir;
            RayInv ray = make_ray(org, rec.dir);
            bool visible;
            if (MODE == kShadowSegment) {
                const RayInv back = make_ray(rec.hp, -rec.dir);
                visible = !trace_occluded<COUNT>(sc, back, 0.1f, rec.sdist - 0.01f, col, kBlock, c);
            } else {
                Hit h;
                if (MODE == kShadowEarly) {
                    // test distance ~ 0.01 + t: a hit this far in front of the shading point is more
                    // than 0.1 from it both ways, with room for rounding in the hit point
                    const float scale =
                        fmaxf(fmaxf(fabsf(org.x), fabsf(org.y)), fmaxf(fabsf(org.z), rec.sdist)) + 1.f;
                    const float t_exit = rec.sdist - 0.11f - 1e-4f * scale;
                    h = trace_closest_wave<COUNT, true, true>(sc, ray, 1e-6f, 1e8f, col, kBlock, c, t_exit);
                } else {
                    h = trace_closest_wave<COUNT>(sc, ray, 1e-6f, 1e8f, col, kBlock, c);
                }
                const bool hit = h.prim >= 0;
                vec3 tp = v3(0.f, 0.f, 0.f);
                if (hit) tp = position_at(sc.tri_attr, h.prim, h.b1, h.b2);
                visible = h.prim != kEarlyExit && shadow_visible(rec, hit, tp);
            }
            if (visible) {
                lit = true;
                int pid = f_as_i(r1.w);
                float4* s = (tag & 1) == 0 ? slot_d : slot_i;
                float4 a = s[pid];
                a.x += r3.x;
                a.y += r3.y;
                a.z += r3.z;
                s[pid] = a;
            }
        }
        if (COUNT) {
            wave_add_u64(&stats[kStatShadowNodes], c.nodes);
            wave_add_u64(&stats[kStatShadowTris], c.tris);
            const uint64_t m = __ballot(lit);
            if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(&stats[kStatShadowLit], (unsigned long long)__popcll(m));
        }
    }
}

// The suspended shadow searches of k_shadow, 64 to a wave: each lane reloads
// its shadow record, puts its stack back and resumes where it stopped.
template <bool COUNT, bool POOLED>
__global__ __launch_bounds__(kBlock) void k_shadow_cont(const JobDev* __restrict__ jobs,
                                                         const float4* __restrict__ shadow, float4* slot_d,
                                                         float4* slot_i, int* counters, unsigned long long* stats,
                                                         ContBufs cont) {
    uint32_t* col = stk_dyn + threadIdx.x;
    SliceCursor cur;
    cur.init(counters, kCntShadowCont, kFetchShadowCont, cont.room);
    int base;
    while (cur.next(base)) {
        const int j = base + (int)(threadIdx.x & (kWave - 1));
        TravCount c{0, 0};
        bool lit = false;
        if (j < cur.count) {
            const size_t k = (size_t)cur.slice * cont.room + j;
            const int4 s = cont.state[k];
            const float4 hh = cont.hit[k];
            const size_t r = (size_t)(uint32_t)s.x;
            for (int e = 0; e < s.w; ++e) col[e * kBlock] = cont.stack[(size_t)e * cont.total + k];
            const float4 r0 = *shadow_field(shadow, r, 0), r1 = *shadow_field(shadow, r, 1);
            const float4 r2 = *shadow_field(shadow, r, 2), r3 = *shadow_field(shadow, r, 3);
            const DevScene& sc = pass_job<POOLED>(jobs, f_as_i(r2.w) >> 1).sc;
            const vec3 dir = v3(r1.x, r1.y, r1.z);
            const RayInv ray = make_ray(v3(r0.x, r0.y, r0.z) + 0.01f * dir, dir);
            TravState st;
            st.h.t = hh.x;
            st.h.b1 = hh.y;
            st.h.b2 = hh.z;
            st.h.prim = f_as_i(hh.w);
            st.ref = (uint32_t)s.z;
            st.sp = s.w;
            trace_closest_run_wave<COUNT>(sc, ray, 1e-6f, col, kBlock, c, st);
            lit = shadow_finish(sc, r0, r1, r2, r3, st.h, slot_d, slot_i);
        }
        if (COUNT) {
            wave_add_u64(&stats[kStatShadowNodes], c.nodes);
            wave_add_u64(&stats[kStatShadowTris], c.tris);
            const uint64_t m = __ballot(lit);
            if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(&stats[kStatShadowLit], (unsigned long long)__popcll(m));
        }
    }
}

// Per-pixel accumulation in sample order, so the sum is the same as adding
// one sample per frame (reference: Src/Core/renderer.cuh:117-122).  Pass
// pixel g is pixel g - pix_base of its job; the job's k samples sit at pass
// paths path_base + s * wh + pixel.
template <bool POOLED>
__global__ __launch_bounds__(kBlock) void k_resolve(const PassHeader* __restrict__ hdr,
                                                     const JobDev* __restrict__ jobs,
                                                     const float4* __restrict__ slot_d,
                                                     const float4* __restrict__ slot_i) {
    const int n_pix = hdr->n_pixels;
    const int n_jobs = POOLED ? hdr->n_jobs : 1;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n_pix; g += gridDim.x * blockDim.x) {
        const JobDev& J = pass_job<POOLED>(jobs, POOLED ? job_of_pixel(jobs, n_jobs, g) : 0);
        const int wh = J.fp.wh, px = g - J.pix_base, k = J.fp.n_paths / wh;
        float4 a = J.acc_d[px], b = J.acc_i[px];
        const size_t base = (size_t)J.path_base + px;
        for (int s = 0; s < k; ++s) {
            float4 x = slot_d[base + (size_t)s * wh];
            float4 y = slot_i[base + (size_t)s * wh];
            a.x += x.x; a.y += x.y; a.z += x.z;
            b.x += y.x; b.y += y.y; b.z += y.z;
        }
        J.acc_d[px] = a;
        J.acc_i[px] = b;
    }
}

__global__ void k_zero4(float4* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Closest hit for an arbitrary ray batch: rays are (o, tmin), (d, tmax).  The
// stack column is dynamic LDS sized to the tree, as in extend.
__global__ __launch_bounds__(kBlock) void k_trace_batch(DevScene sc, const float4* __restrict__ rays, int n,
                                                         float4* __restrict__ out, uint2* __restrict__ counts) {
    uint32_t* col = stk_dyn + threadIdx.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 o = rays[2 * (size_t)i], d = rays[2 * (size_t)i + 1];
        RayInv r = make_ray(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
        TravCount c{0, 0};
        Hit h = trace_closest_wave<true>(sc, r, o.w, d.w, col, kBlock, c);
        out[i] = make_float4(h.t, h.b1, h.b2, i_as_f(h.prim));
        counts[i] = make_uint2(c.nodes, c.tris);
    }
}

// Diffuse-texture lookups at given (u, v): the albedo the shading stage uses.
__global__ __launch_bounds__(kBlock) void k_texture_lookup(DevScene sc, int tex, const float2* __restrict__ uv, int n,
                                                            float4* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 q = uv[i];
    vec3 c = texture_albedo(sc, tex, q.x, q.y);
    out[i] = make_float4(c.x, c.y, c.z, 0.f);
}

// Device math self-test: op 0 div, 1 sqrt, 2 sin, 3 cos, 4 acos, 5 exp, 6 log,
// 7 atan2, 8 philox (a = counter lane bits), 9 inverse, 10 pow.
__global__ void k_math(int op, const float* a, const float* b, float* out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = a[i], y = b[i], r = 0.f;
    switch (op) {
        case 0: r = x / y; break;
        case 1: r = sqrtf(x); break;
        case 2: r = fm::sinf_(x); break;
        case 3: r = fm::cosf_(x); break;
        case 4: r = fm::acosf_(x); break;
        case 5: r = fm::expf_(x); break;
        case 6: r = fm::logf_(x); break;
        case 7: r = fm::atan2f_(x, y); break;
        case 8: {
            uint32_t o[4];
            uint32_t c = (uint32_t)f_as_i(x);
            philox4x32_10(c, c ^ 0x9E3779B9u, (uint32_t)i, 7u, (uint32_t)f_as_i(y), 0x5EEDC0DEu, o);
            r = i_as_f((int32_t)o[i & 3]);
            break;
        }
        case 9: r = 1.f / x; break;
        case 10: r = fm::powf_(x, y); break;
        default: break;
    }
    out[i] = r;
}

// ---------------------------------------------------------------- frames
// Per-pixel frame kernels work on tiles of 2^tile_shift x (256 >> tile_shift)
// pixels (one 256-lane block), so the neighbourhood taps of the filters stay
// within a few cache lines per wave; the shape is a tuning knob (16 x 16 by default).
//
// Tiles go to the XCDs in bands: workgroups are dispatched to the 8 XCDs
// round-robin (XCD = blockIdx % 8), so workgroup b takes tile
// (b % 8) * per + b / 8 with per = ceil(tiles / 8), and XCD x works through the
// x-th eighth of the tiles in raster order -- whole rows of tiles, whose
// neighbours' pixels the filters' taps find in that XCD's own L2 rather than
// in another XCD's.  The grid is rounded up to a multiple of 8; the padding
// workgroups get no tile.
constexpr int kXcds = 8;
inline int banded_grid(int tiles) { return kXcds * ((tiles + kXcds - 1) / kXcds); }
__device__ __forceinline__ int banded_tile(int tiles) {
    const int per = (tiles + kXcds - 1) / kXcds;
    const int t = ((int)blockIdx.x % kXcds) * per + (int)blockIdx.x / kXcds;
    return t < tiles ? t : -1;
}
__device__ __forceinline__ bool tile_pixel(int w, int h, int tile_shift, int& i, int& j, int& tile) {
    const int tw = 1 << tile_shift, th = kBlock >> tile_shift;
    const int tiles_x = (w + tw - 1) >> tile_shift;
    tile = banded_tile(tiles_x * ((h + th - 1) / th));
    if (tile < 0) {
        i = w;
        j = h;
        return false;
    }
    i = (tile % tiles_x) * tw + (threadIdx.x & (tw - 1));
    j = (tile / tiles_x) * th + (threadIdx.x >> tile_shift);
    return i < w && j < h;
}
__device__ __forceinline__ bool tile_pixel(int w, int h, int tile_shift, int& i, int& j) {
    int tile;
    return tile_pixel(w, h, tile_shift, i, j, tile);
}

// The stack column comes from dynamic LDS sized to the tree's depth, as in
// extend.  Tiles in plain raster order, one per workgroup: the G-buffer's cost
// follows the scene's depth complexity, and round-robin spreads the costly
// rows of tiles over all XCDs (in bands, measured 24 % slower).
__global__ __launch_bounds__(kBlock) void k_gbuffer(DevScene sc, GBufferParams g, GBufferView out, int ts) {
    const int tw = 1 << ts, th = kBlock >> ts, tiles_x = (g.width + tw - 1) >> ts;
    const int i = ((int)blockIdx.x % tiles_x) * tw + ((int)threadIdx.x & (tw - 1));
    const int j = ((int)blockIdx.x / tiles_x) * th + ((int)threadIdx.x >> ts);
    if (i >= g.width || j >= g.height) return;
    gbuffer_pixel(sc, g, i, j, stk_dyn + threadIdx.x, kBlock, out);
}

// The temporal pass also lists the tiles that hold a pixel whose history is
// shorter than kVarianceHistory, for the variance pass to visit; the pixels
// with a longer history get the variance pass's output (a copy) from the
// temporal pass itself.  A listed tile is estimated as a whole workgroup, as
// an unlisted pass would: the 7x7 gather is latency-bound, and a list of
// single pixels (measured) leaves too few waves in flight and loses the
// neighbourhood the lanes of a tile share.  One list per XCD band (entries
// band * per + k, per = ceil(tiles / 8)), so the variance pass visits a
// band's tiles on the same XCD; one atomic per workgroup.  Every lane reaches
// the barrier, also those outside the image.
__global__ __launch_bounds__(kBlock) void k_temporal(int w, int h, int ts, int first, SvgfParams p, TemporalIO io,
                                                      GBufferView cur, GBufferView prev, int* __restrict__ tiles_out,
                                                      int* n_tiles) {
    int i, j, tile;
    const bool in = tile_pixel(w, h, ts, i, j, tile);
    int hl = kVarianceHistory;
    if (in) hl = temporal_pixel(i, j, w, h, first != 0, p, io, cur, prev);
    if (__syncthreads_or(in && hl < kVarianceHistory) && threadIdx.x == 0) {
        const int tw = 1 << ts, th = kBlock >> ts;
        const int per = (((w + tw - 1) >> ts) * ((h + th - 1) / th) + kXcds - 1) / kXcds;
        const int band = (int)blockIdx.x % kXcds;
        tiles_out[band * per + atomicAdd(&n_tiles[band], 1)] = tile;
    }
}

// Over the temporal pass's tile lists: workgroup b takes entries b / 8,
// b / 8 + per, ... of band b % 8's list, on the band's XCD; a lane estimates
// its pixel when that pixel's history is short.  The frame after this one
// lists into the other set of counts, which is cleared here (the two passes
// run in stream order, so nothing else touches it meanwhile).
__global__ __launch_bounds__(kBlock) void k_variance(int w, int h, int ts, SvgfParams p, FilterIO io, GBufferView g,
                                                      const int* __restrict__ tiles, const int* n_tiles,
                                                      int* n_tiles_next) {
    if (blockIdx.x < kXcds && threadIdx.x == 0) n_tiles_next[blockIdx.x] = 0;
    const int tw = 1 << ts, th = kBlock >> ts, tiles_x = (w + tw - 1) >> ts;
    const int per = (tiles_x * ((h + th - 1) / th) + kXcds - 1) / kXcds;
    const int band = (int)blockIdx.x % kXcds;
    const int n = n_tiles[band];
    for (int k = (int)blockIdx.x / kXcds; k < n; k += (int)gridDim.x / kXcds) {
        const int t = tiles[band * per + k];
        const int i = (t % tiles_x) * tw + ((int)threadIdx.x & (tw - 1));
        const int j = (t / tiles_x) * th + ((int)threadIdx.x >> ts);
        if (i < w && j < h && io.len[(size_t)j * w + i] < kVarianceHistory) variance_estimate(i, j, w, h, p, io, g);
    }
}

__global__ __launch_bounds__(kBlock) void k_atrous(int w, int h, int ts, int step, SvgfParams p, FilterIO io,
                                                    GBufferView g) {
    int i, j;
    if (!tile_pixel(w, h, ts, i, j)) return;
    atrous_pixel(i, j, w, h, step, p, io, g);
}

// LDS-tiled first a-trous pass (step 1).  A block takes a 16x16 pixel tile,
// stages the tile plus a 2-pixel apron -- guide key/length, guide
// normal/depth and both input colours, 56 B per cell -- in LDS with one
// global load per cell and array, and reads every tap (and the 3x3 variance
// neighbourhood) from there; cells outside the image carry history length
// -1.  Same arithmetic in the same order as the untiled pass (atrous_core).
// Measured per pass on 1280x720: 152 us against 165 us untiled.  The same
// tiling on the step-s lattice (pixels with equal coordinates modulo s) tied
// at steps 2 and 4 and lost at 8 and 16 (186 and 264 us), where filling the
// tile is a strided gather; those passes stay untiled.
constexpr int kAtTile = 16, kAtApron = 2, kAtSide = kAtTile + 2 * kAtApron;
static_assert(kAtTile * kAtTile == kBlock, "one output pixel per thread");

struct AtrousTile {
    const int2* kl;
    const float4 *nd, *cd, *ci;
    int lx, ly;
    __device__ int cell(int di, int dj) const { return (ly + kAtApron + dj) * kAtSide + (lx + kAtApron + di); }
    __device__ bool key(int ki, int kj, int2& k) const {
        k = kl[cell(ki, kj)];
        return k.y >= 0;
    }
    __device__ void data(int ki, int kj, float4& n, float4& d, float4& c) const {
        const int q = cell(ki, kj);
        n = nd[q];
        d = cd[q];
        c = ci[q];
    }
    __device__ bool var(int vi, int vj, float& vd, float& vv) const {
        const int q = cell(vi, vj);
        if (kl[q].y < 0) return false;
        vd = cd[q].w;
        vv = ci[q].w;
        return true;
    }
};

__global__ __launch_bounds__(kBlock) void k_atrous_tile1(int w, int h, int tiles_x, int tiles_y, SvgfParams p,
                                                          FilterIO io, GBufferView g) {
    __shared__ int2 s_kl[kAtSide * kAtSide];
    __shared__ float4 s_nd[kAtSide * kAtSide], s_cd[kAtSide * kAtSide], s_ci[kAtSide * kAtSide];
    const int tile = banded_tile(tiles_x * tiles_y);
    if (tile < 0) return;
    const int x0 = (tile % tiles_x) * kAtTile, y0 = (tile / tiles_x) * kAtTile;
    for (int k = threadIdx.x; k < kAtSide * kAtSide; k += kBlock) {
        const int x = x0 - kAtApron + k % kAtSide, y = y0 - kAtApron + k / kAtSide;
        if (x >= 0 && x < w && y >= 0 && y < h) {
            const size_t n = (size_t)y * w + x;
            s_kl[k] = io.guide_kl[n];
            s_nd[k] = io.guide_nd[n];
            s_cd[k] = io.in_d[n];
            s_ci[k] = io.in_i[n];
        } else {
            s_kl[k] = make_int2(0, -1);
        }
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % kAtTile, ly = (int)threadIdx.x / kAtTile;
    const int i = x0 + lx, j = y0 + ly;
    if (i >= w || j >= h) return;
    atrous_core((size_t)j * w + i, 1, p, io, g, AtrousTile{s_kl, s_nd, s_cd, s_ci, lx, ly});
}

// Denoised colour = filtered direct + indirect (GatherKernel's denoise branch,
// reference: renderer.cuh:970-978, before the display gamma).
__global__ __launch_bounds__(kBlock) void k_combine(const float4* __restrict__ d, const float4* __restrict__ ind,
                                                     int n, float4* out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        float4 a = d[k], b = ind[k];
        out[k] = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, 1.f);
    }
}

// ======================================================================
// Host side
// ======================================================================

// Triangle k of the BVH order (original triangle t = order[k]): the traversal
// record (v0, e1 = v1 - v0, e2 = v2 - v0) and the 128-byte attribute line.
// mesh_tab holds (first triangle, material, light, shading class) per mesh.
__global__ void k_pack_triangles(int n, const int32_t* __restrict__ order, const float* __restrict__ pos,
                                 const float* __restrict__ nrm, const float* __restrict__ uv,
                                 const float* __restrict__ gn, const int32_t* __restrict__ mesh,
                                 const int4* __restrict__ mesh_tab, float4* __restrict__ isect,
                                 float4* __restrict__ attr) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int t = order[k];
    const float* p = pos + 9 * (size_t)t;
    const float* nn = nrm + 9 * (size_t)t;
    const float* w = uv + 6 * (size_t)t;
    const float* g = gn + 3 * (size_t)t;
    const int m = mesh[t];
    const int4 mr = mesh_tab[m];
    isect[3 * (size_t)k + 0] = make_float4(p[0], p[1], p[2], 0.f);
    isect[3 * (size_t)k + 1] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], 0.f);
    isect[3 * (size_t)k + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.f);
    float4* a = attr + kAttrStride * (size_t)k;
    a[0] = make_float4(p[0], p[1], p[2], i_as_f(t - mr.x));
    a[1] = make_float4(p[3], p[4], p[5], g[0]);
    a[2] = make_float4(p[6], p[7], p[8], g[1]);
    a[3] = make_float4(nn[0], nn[1], nn[2], g[2]);
    a[4] = make_float4(nn[3], nn[4], nn[5], w[0]);
    a[5] = make_float4(nn[6], nn[7], nn[8], w[1]);
    a[6] = make_float4(w[2], w[3], w[4], w[5]);
    a[7] = make_float4(i_as_f(m), i_as_f(mr.y), i_as_f(mr.z), i_as_f(mr.w));
}

// ------------------------------------------------------------------ refit
// New positions for the same triangles over the same tree.  Box unions keep
// std::min / std::max tie rules (the first operand wins a tie, which decides
// between -0 and +0) so the device refit writes the host refit's bits.
__device__ __forceinline__ float rf_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float rf_max(float a, float b) { return (a < b) ? b : a; }

// Position-dependent words of the triangle records: the intersection record
// and the vertices and geometric normal of the attribute line (normalised
// cross product, as the triangle constructor on the host forms it); shading
// normals too when `nrm` is given.  Indices, uvs and ids are kept.
__global__ void k_refit_triangles(int n, const int32_t* __restrict__ order, const float* __restrict__ pos,
                                  const float* __restrict__ nrm, float4* __restrict__ isect,
                                  float4* __restrict__ attr) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int t = order[k];
    const float* p = pos + 9 * (size_t)t;
    const vec3 a = v3(p[0], p[1], p[2]), b = v3(p[3], p[4], p[5]), c = v3(p[6], p[7], p[8]);
    const vec3 g = normalize(cross(b - a, c - a));
    isect[3 * (size_t)k + 0] = make_float4(p[0], p[1], p[2], 0.f);
    isect[3 * (size_t)k + 1] = make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], 0.f);
    isect[3 * (size_t)k + 2] = make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.f);
    float4* A = attr + kAttrStride * (size_t)k;
    const float4 a0 = A[0], a3 = A[3];
    A[0] = make_float4(p[0], p[1], p[2], a0.w);
    A[1] = make_float4(p[3], p[4], p[5], g.x);
    A[2] = make_float4(p[6], p[7], p[8], g.y);
    if (nrm) {
        const float* q = nrm + 9 * (size_t)t;
        const float4 a4 = A[4], a5 = A[5];
        A[3] = make_float4(q[0], q[1], q[2], g.z);
        A[4] = make_float4(q[3], q[4], q[5], a4.w);
        A[5] = make_float4(q[6], q[7], q[8], a5.w);
    } else {
        A[3] = make_float4(a3.x, a3.y, a3.z, g.z);
    }
}

// Parent link of every interior node but the root: (parent << 1) | side.
__global__ void k_refit_parents(int ni, const PackedNode* __restrict__ nodes, int* __restrict__ parent) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ni) return;
    const int4 w3 = nodes[p].w3;
    if (!((uint32_t)w3.x & kLeafBit)) parent[w3.x] = p << 1;
    if (!((uint32_t)w3.y & kLeafBit)) parent[w3.y] = (p << 1) | 1;
}

// A child box is 6 consecutive floats of its parent's line: side 0 = words
// 0..5 (left lo.xyz, hi.xyz), side 1 = words 6..11.  Slots are written and
// read by different workgroups within one launch, so they go through
// device-scope atomic loads and stores ordered by fences.
__device__ __forceinline__ void slot_put(PackedNode* nd, int side, const float* b) {
    uint32_t* u = reinterpret_cast<uint32_t*>(nd) + 6 * side;
    for (int i = 0; i < 6; ++i) __hip_atomic_store(u + i, __float_as_uint(b[i]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void slot_get(PackedNode* nd, int side, float* b) {
    uint32_t* u = reinterpret_cast<uint32_t*>(nd) + 6 * side;
    for (int i = 0; i < 6; ++i) b[i] = __uint_as_float(__hip_atomic_load(u + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Bottom-up box refit, one thread per interior node.  A thread writes its
// node's leaf-child boxes (union of the leaf's triangle boxes in leaf order,
// from the empty box) and adds their number to the node's counter; whoever
// brings a counter to 2 owns the node: it forms the node's box (left child's
// box expanded by the right's), writes it into the parent's slot and adds 1
// to the parent's counter, and so on up to the root.  No thread waits on
// another, so every thread ends after at most depth steps.
__global__ void k_refit_nodes(int ni, PackedNode* nodes, const int* __restrict__ parent, int* cnt,
                              const int32_t* __restrict__ order, const float* __restrict__ pos) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ni) return;
    const int4 w3 = nodes[p].w3;  // child references: the refit never writes them
    const float fmax_ = (float)3.40282e+038;
    int leaves = 0;
    for (int side = 0; side < 2; ++side) {
        const uint32_t ref = (uint32_t)(side ? w3.y : w3.x);
        if (!(ref & kLeafBit)) continue;
        const uint32_t first = (ref & ~kLeafBit) >> 3, count = ref & 7u;
        float b[6] = {fmax_, fmax_, fmax_, -fmax_, -fmax_, -fmax_};
        for (uint32_t k = 0; k < count; ++k) {
            const float* q = pos + 9 * (size_t)order[first + k];
            for (int a = 0; a < 3; ++a) {
                b[a] = rf_min(b[a], rf_min(q[a], rf_min(q[3 + a], q[6 + a])));
                b[3 + a] = rf_max(b[3 + a], rf_max(q[a], rf_max(q[3 + a], q[6 + a])));
            }
        }
        slot_put(nodes + p, side, b);
        ++leaves;
    }
    if (leaves == 0) return;
    __threadfence();
    if (atomicAdd(&cnt[p], leaves) + leaves < 2) return;
    int node = p;
    while (true) {
        __threadfence();
        float l[6], r[6], u[6];
        slot_get(nodes + node, 0, l);
        slot_get(nodes + node, 1, r);
        for (int a = 0; a < 3; ++a) {
            u[a] = rf_min(l[a], r[a]);
            u[3 + a] = rf_max(l[3 + a], r[3 + a]);
        }
        const int pr = parent[node];
        if (pr < 0) return;  // the root: its box is read by the host from its two slots
        slot_put(nodes + (pr >> 1), pr & 1, u);
        __threadfence();
        if (atomicAdd(&cnt[pr >> 1], 1) == 0) return;  // the sibling is not done yet
        node = pr >> 1;
    }
}

struct GpuScene::Impl {
    DevBuf nodes, tri_isect, tri_attr, materials, lights, light_tris, env_rgb, env_row, env_marg, env_alias_row,
        env_alias_marg, tex_texels, tex_info;
    DevBuf order;             // BVH slot -> scene triangle, kept for refits
    DevBuf parent, refit_cnt; // per interior node, allocated by the first refit
    DevBuf media_root, media_nodes, media_isect, media_prim, media_grid, media_clear;  // random-walk media
    int num_interior = 0;     // packed nodes of the tree (0: the root is a leaf)
    DevScene dev{};
    int max_stack = 1;
    bool has_class[kNumClasses] = {false, false, false, false, false};
    int device = 0;
};

static void upload(DevBuf& b, const void* src, size_t bytes) {
    b.alloc(bytes);
    if (bytes) HIP_OK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

// The media trees packed back to back: node references made absolute (an
// interior index plus the tree's first node, a leaf's first triangle plus the
// tree's first triangle), triangle records in each tree's order with the
// scene's FP32 edge subtraction, and each record's position in the scene's
// BVH order.  One dummy entry per array when the scene has no medium.
struct MediaTables {
    std::vector<uint32_t> root;
    std::vector<PackedNode> nodes;
    std::vector<float4> isect;
    std::vector<int32_t> prim;
    std::vector<float4> grid;   // 2 per mesh
    std::vector<float> clear;
    int max_stack = 1;
};

static void media_tables(const HostScene& sc, MediaTables& t) {
    t.root.assign(std::max<size_t>(sc.meshes.size(), 1), kNoMedium);
    t.grid.assign(2 * t.root.size(), make_float4(0.f, 0.f, 0.f, 0.f));
    t.nodes.clear();
    t.isect.clear();
    t.prim.clear();
    t.clear.clear();
    t.max_stack = 1;
    std::vector<int32_t> slot;  // scene triangle -> position in the scene's BVH order
    if (!sc.media.empty()) {
        slot.resize(sc.bvh.order.size());
        for (size_t k = 0; k < sc.bvh.order.size(); ++k) slot[sc.bvh.order[k]] = (int32_t)k;
    }
    for (const MediumTree& mt : sc.media) {
        const BvhBuild& b = mt.bvh;
        const MeshRec& mr = sc.meshes[mt.mesh];
        const uint32_t node0 = (uint32_t)t.nodes.size(), tri0 = (uint32_t)(t.isect.size() / 3);
        if ((size_t)tri0 + b.order.size() >= ((size_t)1 << 28)) throw std::runtime_error("media exceed 2^28 triangles");
        auto shift = [&](uint32_t ref) {
            if (ref & kLeafBit) return leaf_ref(((ref & ~kLeafBit) >> 3) + tri0, ref & 7u);
            return ref + node0;
        };
        for (PackedNode n : b.packed) {
            n.w3.x = (int)shift((uint32_t)n.w3.x);
            n.w3.y = (int)shift((uint32_t)n.w3.y);
            t.nodes.push_back(n);
        }
        for (int32_t local : b.order) {
            const int tri = mr.start + local;
            const float* p = sc.pos.data() + 9 * (size_t)tri;
            t.isect.push_back(make_float4(p[0], p[1], p[2], 0.f));
            t.isect.push_back(make_float4(p[3] - p[0], p[4] - p[1], p[5] - p[2], 0.f));
            t.isect.push_back(make_float4(p[6] - p[0], p[7] - p[1], p[8] - p[2], 0.f));
            t.prim.push_back(slot[tri]);
        }
        t.root[mt.mesh] = shift(b.root_ref);
        t.max_stack = std::max(t.max_stack, b.max_stack);
        t.grid[2 * mt.mesh] = make_float4(mt.grid_lo.x, mt.grid_lo.y, mt.grid_lo.z, i_as_f(mt.grid_n));
        t.grid[2 * mt.mesh + 1] = make_float4(mt.grid_inv.x, mt.grid_inv.y, mt.grid_inv.z, i_as_f((int)t.clear.size()));
        t.clear.insert(t.clear.end(), mt.clearance.begin(), mt.clearance.end());
    }
    if (t.clear.empty()) t.clear.resize(1);
    if (t.nodes.empty()) t.nodes.resize(1);
    if (t.isect.empty()) t.isect.resize(3);
    if (t.prim.empty()) t.prim.resize(1);
}

// Area-light records and their triangles (v0, v1, v2, geometric normal) in
// light order; one dummy entry each when there are none, so pointers are valid.
static void light_tables(const HostScene& sc, std::vector<LightRec>& lights, std::vector<float4>& ltris) {
    lights = sc.lights;
    ltris.clear();
    for (const auto& L : sc.lights) {
        if (L.type != kAreaLight || L.mesh < 0) continue;
        const MeshRec& m = sc.meshes[L.mesh];
        for (int t = m.start; t < m.start + m.count; ++t) {
            const float* p = &sc.pos[9 * (size_t)t];
            const float* g = &sc.gnormal[3 * (size_t)t];
            ltris.push_back(make_float4(p[0], p[1], p[2], 0.f));
            ltris.push_back(make_float4(p[3], p[4], p[5], 0.f));
            ltris.push_back(make_float4(p[6], p[7], p[8], 0.f));
            ltris.push_back(make_float4(g[0], g[1], g[2], 0.f));
        }
    }
    if (lights.empty()) lights.resize(1);
    if (ltris.empty()) ltris.resize(4);
}

GpuScene::GpuScene(const HostScene& sc) : impl_(new Impl) {
    Impl& I = *impl_;
    HIP_OK(hipGetDevice(&I.device));
    const BvhBuild& bvh = sc.bvh;
    int n = sc.num_tris();
    if (bvh.max_stack > 64) throw std::runtime_error("BVH deeper than the 64-entry traversal stack");
    I.max_stack = std::max(1, bvh.max_stack);
    // Packed nodes; a leaf-only tree gets one dummy line so the pointer is valid.
    std::vector<PackedNode> nodes = bvh.packed;
    I.num_interior = (int)nodes.size();
    if (nodes.empty()) nodes.resize(1);
    upload(I.nodes, nodes.data(), nodes.size() * sizeof(PackedNode));
    // Triangle lines in BVH order, packed on the GPU from the scene-order arrays.
    auto class_of = [&](int material) -> int {
        switch (sc.materials[material].type) {
            case kLight: return kClsTerminal;
            case kPbr: return kClsPbr;
            case kDielectric: return kClsDielectric;
            case kSubsurface: return kClsSubsurface;
            default: return kClsDiffuse;
        }
    };
    std::vector<int4> mesh_tab(std::max<size_t>(sc.meshes.size(), 1));
    for (size_t m = 0; m < sc.meshes.size(); ++m) {
        const MeshRec& mr = sc.meshes[m];
        int cls = class_of(mr.material);
        if (mr.count > 0) I.has_class[cls] = true;
        mesh_tab[m] = make_int4(mr.start, mr.material, mr.light, cls);
    }
    {
        DevBuf dpos, dnrm, duv, dgn, dmesh, dtab;
        upload(dpos, sc.pos.data(), sc.pos.size() * sizeof(float));
        upload(dnrm, sc.nrm.data(), sc.nrm.size() * sizeof(float));
        upload(duv, sc.uv.data(), sc.uv.size() * sizeof(float));
        upload(dgn, sc.gnormal.data(), sc.gnormal.size() * sizeof(float));
        upload(dmesh, sc.mesh.data(), sc.mesh.size() * sizeof(int32_t));
        upload(I.order, bvh.order.data(), bvh.order.size() * sizeof(int32_t));
        upload(dtab, mesh_tab.data(), mesh_tab.size() * sizeof(int4));
        I.tri_isect.alloc(3 * (size_t)n * sizeof(float4));
        I.tri_attr.alloc(kAttrStride * (size_t)n * sizeof(float4));
        if (n > 0)
            k_pack_triangles<<<(n + 255) / 256, 256>>>(n, I.order.as<int32_t>(), dpos.as<float>(), dnrm.as<float>(),
                                                   duv.as<float>(), dgn.as<float>(), dmesh.as<int32_t>(),
                                                   dtab.as<int4>(), I.tri_isect.as<float4>(), I.tri_attr.as<float4>());
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
    }
    std::vector<MaterialRec> mats = sc.materials;
    if (mats.empty()) mats.resize(1);
    upload(I.materials, mats.data(), mats.size() * sizeof(MaterialRec));
    MediaTables media;
    media_tables(sc, media);
    if (media.max_stack > 64) throw std::runtime_error("medium tree deeper than the 64-entry traversal stack");
    I.max_stack = std::max(I.max_stack, media.max_stack);
    upload(I.media_root, media.root.data(), media.root.size() * sizeof(uint32_t));
    upload(I.media_nodes, media.nodes.data(), media.nodes.size() * sizeof(PackedNode));
    upload(I.media_isect, media.isect.data(), media.isect.size() * sizeof(float4));
    upload(I.media_prim, media.prim.data(), media.prim.size() * sizeof(int32_t));
    upload(I.media_grid, media.grid.data(), media.grid.size() * sizeof(float4));
    upload(I.media_clear, media.clear.data(), media.clear.size() * sizeof(float));
    std::vector<LightRec> lights;
    std::vector<float4> ltris;
    light_tables(sc, lights, ltris);
    upload(I.lights, lights.data(), lights.size() * sizeof(LightRec));
    upload(I.light_tris, ltris.data(), ltris.size() * sizeof(float4));

    DevScene& d = I.dev;
    d.nodes = I.nodes.as<PackedNode>();
    d.root_ref = bvh.root_ref;
    d.tri_isect = I.tri_isect.as<float4>();
    d.tri_attr = I.tri_attr.as<float4>();
    d.materials = I.materials.as<MaterialRec>();
    d.media_root = I.media_root.as<uint32_t>();
    d.media_nodes = I.media_nodes.as<PackedNode>();
    d.media_isect = I.media_isect.as<float4>();
    d.media_prim = I.media_prim.as<int32_t>();
    d.media_grid = I.media_grid.as<float4>();
    d.media_clear = I.media_clear.as<float>();
    d.lights = I.lights.as<LightRec>();
    d.num_lights = (int)sc.lights.size();
    d.env_light = sc.env_light;
    d.light_tris = I.light_tris.as<float4>();
    d.bounds_lo = bvh.bounds.lo;
    d.bounds_hi = bvh.bounds.hi;
    d.env_rgb = nullptr;
    d.env_w = d.env_h = 0;
    d.env_row_cdf = d.env_marg_cdf = d.env_pdf = nullptr;
    d.env_alias_row = d.env_alias_marg = nullptr;
    d.env_alias = 0;
    if (sc.env_light >= 0) {
        EnvTables t;
        build_env_tables(sc.env, t);
        upload(I.env_rgb, sc.env.rgb.data(), sc.env.rgb.size() * sizeof(float));
        upload(I.env_row, t.row_cdf.data(), t.row_cdf.size() * sizeof(float));
        upload(I.env_marg, t.marg_cdf.data(), t.marg_cdf.size() * sizeof(float));
        d.env_rgb = I.env_rgb.as<float>();
        d.env_w = sc.env.width;
        d.env_h = sc.env.height;
        d.env_row_cdf = I.env_row.as<float>();
        d.env_marg_cdf = I.env_marg.as<float>();
        upload(I.env_alias_row, t.alias_row.data(), t.alias_row.size() * sizeof(float));
        upload(I.env_alias_marg, t.alias_marg.data(), t.alias_marg.size() * sizeof(float));
        d.env_alias_row = I.env_alias_row.as<float2>();
        d.env_alias_marg = I.env_alias_marg.as<float2>();
    }
    // Diffuse textures, back to back; an empty slot keeps a (0, 0, 0) entry.
    d.tex_texels = nullptr;
    d.tex_info = nullptr;
    if (!sc.textures.empty()) {
        std::vector<int32_t> info(4 * sc.textures.size(), 0);
        size_t texels = 0;
        for (size_t i = 0; i < sc.textures.size(); ++i) {
            const LdrImage& t = sc.textures[i];
            info[4 * i] = (int32_t)texels;
            info[4 * i + 1] = t.w;
            info[4 * i + 2] = t.h;
            texels += (size_t)t.w * t.h;
        }
        if (texels > (size_t)INT32_MAX) throw std::runtime_error("textures exceed 2^31 texels");
        std::vector<uint8_t> all;
        all.reserve(4 * std::max<size_t>(texels, 1));
        for (const LdrImage& t : sc.textures) all.insert(all.end(), t.rgba.begin(), t.rgba.end());
        if (all.empty()) all.resize(4);
        upload(I.tex_texels, all.data(), all.size());
        upload(I.tex_info, info.data(), info.size() * sizeof(int32_t));
        d.tex_texels = I.tex_texels.as<uint32_t>();
        d.tex_info = I.tex_info.as<int32_t>();
    }
}

GpuScene::~GpuScene() { delete impl_; }

// The box a node's two child slots add up to: the left child's box expanded
// by the right child's, with the device refit's tie rules.
static Box union_of_slots(const PackedNode& p) {
    auto mn = [](float a, float b) { return (b < a) ? b : a; };
    auto mx = [](float a, float b) { return (a < b) ? b : a; };
    Box b;
    b.lo = v3(mn(p.w0.x, p.w1.z), mn(p.w0.y, p.w1.w), mn(p.w0.z, p.w2.x));
    b.hi = v3(mx(p.w0.w, p.w2.y), mx(p.w1.x, p.w2.z), mx(p.w1.y, p.w2.w));
    return b;
}

void GpuScene::refit(HostScene& host, const float* nrm, double phase_ms[2]) {
    Impl& I = *impl_;
    const int n = host.num_tris();
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&]() {
        auto t = std::chrono::steady_clock::now();
        double ms = std::chrono::duration<double, std::milli>(t - t0).count();
        t0 = t;
        return ms;
    };
    HIP_OK(hipSetDevice(I.device));
    HIP_OK(hipDeviceSynchronize());  // renders in flight read the arrays rewritten below
    DevBuf dpos, dnrm;
    upload(dpos, host.pos.data(), host.pos.size() * sizeof(float));
    if (nrm) upload(dnrm, host.nrm.data(), host.nrm.size() * sizeof(float));
    std::vector<LightRec> lights;
    std::vector<float4> ltris;
    light_tables(host, lights, ltris);
    if (lights.size() * sizeof(LightRec) != I.lights.bytes || ltris.size() * sizeof(float4) != I.light_tris.bytes)
        throw std::runtime_error("refit: light tables changed size");
    HIP_OK(hipMemcpy(I.lights.p, lights.data(), I.lights.bytes, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(I.light_tris.p, ltris.data(), I.light_tris.bytes, hipMemcpyHostToDevice));
    phase_ms[0] = lap();
    if (n > 0)
        k_refit_triangles<<<(n + 255) / 256, 256>>>(n, I.order.as<int32_t>(), dpos.as<float>(),
                                                    nrm ? dnrm.as<float>() : nullptr, I.tri_isect.as<float4>(),
                                                    I.tri_attr.as<float4>());
    HIP_OK(hipGetLastError());
    const int ni = I.num_interior;
    Box root;
    if (ni > 0) {
        if (!I.parent.p) {
            I.parent.alloc((size_t)ni * sizeof(int));
            I.refit_cnt.alloc((size_t)ni * sizeof(int));
            HIP_OK(hipMemset(I.parent.p, 0xff, I.parent.bytes));  // the root keeps -1
            k_refit_parents<<<(ni + 255) / 256, 256>>>(ni, I.nodes.as<PackedNode>(), I.parent.as<int>());
            HIP_OK(hipGetLastError());
        }
        HIP_OK(hipMemset(I.refit_cnt.p, 0, I.refit_cnt.bytes));
        k_refit_nodes<<<(ni + 255) / 256, 256>>>(ni, I.nodes.as<PackedNode>(), I.parent.as<int>(),
                                                 I.refit_cnt.as<int>(), I.order.as<int32_t>(), dpos.as<float>());
        HIP_OK(hipGetLastError());
        PackedNode top;
        HIP_OK(hipMemcpy(&top, I.nodes.p, sizeof(PackedNode), hipMemcpyDeviceToHost));  // waits for the kernels
        root = union_of_slots(top);
    } else {
        HIP_OK(hipDeviceSynchronize());
        refit_bvh(host.pos.data(), n, host.bvh);  // the whole tree is one leaf
        root = host.bvh.bounds;
    }
    host.bvh.bounds = root;
    I.dev.bounds_lo = root.lo;
    I.dev.bounds_hi = root.hi;
    // Media trees: refit on the host (the caller refit them), rewritten in place.
    if (!host.media.empty()) {
        MediaTables media;
        media_tables(host, media);
        if (media.nodes.size() * sizeof(PackedNode) != I.media_nodes.bytes ||
            media.isect.size() * sizeof(float4) != I.media_isect.bytes)
            throw std::runtime_error("refit: media trees changed size");
        HIP_OK(hipMemcpy(I.media_nodes.p, media.nodes.data(), I.media_nodes.bytes, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(I.media_isect.p, media.isect.data(), I.media_isect.bytes, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(I.media_grid.p, media.grid.data(), I.media_grid.bytes, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(I.media_clear.p, media.clear.data(), I.media_clear.bytes, hipMemcpyHostToDevice));
    }
    phase_ms[1] = lap();
}

void GpuScene::read_bvh(BvhBuild& b) const {
    const Impl& I = *impl_;
    if (I.num_interior == 0) return;  // a one-leaf tree is refit on the host
    if ((int)b.packed.size() != I.num_interior) throw std::runtime_error("read_bvh: node count differs");
    HIP_OK(hipSetDevice(I.device));
    HIP_OK(hipMemcpy(b.packed.data(), I.nodes.p, b.packed.size() * sizeof(PackedNode), hipMemcpyDeviceToHost));
    unpack_bvh(b.packed, b.root_ref, b.bounds, b.nodes);
}

const DevScene& GpuScene::dev() const { return impl_->dev; }
int GpuScene::max_stack() const { return impl_->max_stack; }
bool GpuScene::has_class(int c) const { return impl_->has_class[c]; }
size_t GpuScene::device_bytes() const {
    const Impl& I = *impl_;
    return I.nodes.bytes + I.tri_isect.bytes + I.tri_attr.bytes + I.materials.bytes + I.lights.bytes +
           I.light_tris.bytes + I.env_rgb.bytes + I.env_row.bytes + I.env_marg.bytes + I.env_alias_row.bytes +
           I.env_alias_marg.bytes + I.tex_texels.bytes + I.tex_info.bytes + I.order.bytes + I.parent.bytes +
           I.refit_cnt.bytes + I.media_root.bytes + I.media_nodes.bytes + I.media_isect.bytes + I.media_prim.bytes +
           I.media_grid.bytes + I.media_clear.bytes;
}

// Row-conditional and marginal CDFs weighted by sin(theta) (reference:
// Src/Core/texture.cu:38-80).
void build_alias_table(const double* w, int n, float* prob, int32_t* alias) {
    double total = 0.0;
    for (int i = 0; i < n; ++i) total += w[i];
    if (!(total > 0.0)) {
        for (int i = 0; i < n; ++i) {
            prob[i] = 1.f;
            alias[i] = i;
        }
        return;
    }
    std::vector<double> q(n);
    std::vector<int> small, large;
    small.reserve(n);
    large.reserve(n);
    for (int i = 0; i < n; ++i) {
        q[i] = (w[i] * (double)n) / total;
        if (q[i] < 1.0) small.push_back(i);
        else large.push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int a = small.back(), b = large.back();
        small.pop_back();
        large.pop_back();
        prob[a] = (float)q[a];
        alias[a] = b;
        q[b] = (q[b] + q[a]) - 1.0;
        if (q[b] < 1.0) small.push_back(b);
        else large.push_back(b);
    }
    while (!large.empty()) {
        prob[large.back()] = 1.f;
        alias[large.back()] = large.back();
        large.pop_back();
    }
    while (!small.empty()) {
        prob[small.back()] = 1.f;
        alias[small.back()] = small.back();
        small.pop_back();
    }
}

void build_env_tables(const EnvMap& env, EnvTables& t) {
    int w = env.width, h = env.height;
    t.row_cdf.assign((size_t)w * h, 0.f);
    t.marg_cdf.assign(h, 0.f);
    for (int y = 0; y < h; ++y) {
        float vsum = 0.0f;
        float theta = (float)((double)(((float)y + 0.5f) / (float)h) * 3.1415926);
        float st = sinf(theta);
        for (int x = 0; x < w; ++x) {
            size_t idx = (size_t)y * w + x;
            const float* c = &env.rgb[3 * idx];
            float lum = (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2];
            lum *= st;
            vsum += lum;
            t.row_cdf[idx] = vsum;
        }
        t.marg_cdf[y] = vsum;
    }
    for (int y = 0; y < h; ++y) {
        float rs = t.marg_cdf[y];
        for (int x = 0; x < w; ++x) {
            size_t idx = (size_t)y * w + x;
            t.row_cdf[idx] = rs > 0.0f ? t.row_cdf[idx] / rs : 0.0f;
        }
    }
    // Alias tables over the same weights: texel luminance * sin(theta) per
    // row, the row totals for the marginal.
    {
        t.alias_row.assign(2 * (size_t)w * h, 0.f);
        t.alias_marg.assign(2 * (size_t)h, 0.f);
        std::vector<double> wt(std::max(w, h));
        std::vector<float> pr(std::max(w, h));
        std::vector<int32_t> al(std::max(w, h));
        for (int y = 0; y < h; ++y) {
            float theta = (float)((double)(((float)y + 0.5f) / (float)h) * 3.1415926);
            float st = sinf(theta);
            for (int x = 0; x < w; ++x) {
                const float* c = &env.rgb[3 * ((size_t)y * w + x)];
                float lum = (0.2126f * c[0] + 0.7152f * c[1]) + 0.0722f * c[2];
                wt[x] = (double)(lum * st);
            }
            build_alias_table(wt.data(), w, pr.data(), al.data());
            for (int x = 0; x < w; ++x) {
                t.alias_row[2 * ((size_t)y * w + x)] = pr[x];
                std::memcpy(&t.alias_row[2 * ((size_t)y * w + x) + 1], &al[x], 4);
            }
        }
        for (int y = 0; y < h; ++y) wt[y] = (double)t.marg_cdf[y];
        build_alias_table(wt.data(), h, pr.data(), al.data());
        for (int y = 0; y < h; ++y) {
            t.alias_marg[2 * (size_t)y] = pr[y];
            std::memcpy(&t.alias_marg[2 * (size_t)y + 1], &al[y], 4);
        }
    }
    for (int i = 1; i < h; ++i) t.marg_cdf[i] += t.marg_cdf[i - 1];
    float total = h > 0 ? t.marg_cdf[h - 1] : 0.f;
    for (int i = 0; i < h; ++i) t.marg_cdf[i] = total > 0.0f ? t.marg_cdf[i] / total : 0.0f;
}

// ------------------------------------------------------------------ profiler
hipEvent_t Profiler::begin(hipStream_t s) {
    if (!on) return nullptr;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, s));
    return e;
}

void Profiler::end(int stage, hipEvent_t b, hipStream_t s) {
    if (!on) return;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, s));
    ev.push_back({stage, {b, e}});
}

void Profiler::collect() {
    for (auto& e : ev) {
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, e.second.first, e.second.second));
        ms[e.first] += t;
        launches[e.first] += 1;
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    ev.clear();
}

void Profiler::reset() {
    for (auto& e : ev) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    ev.clear();
    for (int s = 0; s < kNumStages; ++s) {
        ms[s] = 0;
        launches[s] = 0;
    }
}

Profiler::~Profiler() {
    for (auto& e : ev) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
}

// ------------------------------------------------------------------ engine
template <typename K>
static int persistent_grid(K kernel, int cus, size_t lds = 0) {
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, lds));
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

Engine::Engine(size_t capacity_paths, bool profile) {
    prof.on = profile;
    if (capacity_paths < 1) capacity_paths = 1;
    if (capacity_paths > (size_t)0x7fffffff) capacity_paths = 0x7fffffff;
    cap_ = capacity_paths;
    HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    int dev = 0, cus = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int k = 0; k < 2; ++k) {
        qo_[k].alloc(cap_ * 16);
        qd_[k].alloc(cap_ * 16);
        qw_[k].alloc(cap_ * 16);
    }
    hits_.alloc(cap_ * 16);
    lists_.alloc(cap_ * 4 * kNumClasses);
    shadow_.alloc(((cap_ + 63) & ~(size_t)63) * 64);
    slot_d_.alloc(cap_ * 16);
    slot_i_.alloc(cap_ * 16);
    cont_room_ = (int)std::max<size_t>(1, cap_ / (kContPathsPer * kSlices));
    const size_t cont_total = (size_t)cont_room_ * kSlices;
    cont_hit_.alloc(cont_total * 16);
    cont_state_.alloc(cont_total * 16);
    cont_stack_.alloc(cont_total * 4 * kContStack);
    counters_.alloc(kCounterInts * sizeof(int));
    stats_.alloc(kNumStats * sizeof(unsigned long long));
    tables_.alloc(kPassTableBytes);
    HIP_OK(hipHostMalloc(&host_tables_, kPassTableBytes * kRing, hipHostMallocDefault));
    for (int k = 0; k < kRing; ++k) HIP_OK(hipEventCreateWithFlags(&ring_ev_[k], hipEventDisableTiming));
    HIP_OK(hipMemset(counters_.p, 0, kCounterInts * sizeof(int)));
    HIP_OK(hipMemset(stats_.p, 0, kNumStats * sizeof(unsigned long long)));
    cus_ = cus;
    grid_linear_ = cus * 8;
    HIP_OK(hipDeviceSynchronize());
}

void Engine::set_ray_sort(bool on) {
    if (on && !keys_.p) {
        keys_.alloc(cap_ * sizeof(uint16_t));
        perm_.alloc(cap_ * sizeof(int));
        sorted_o_.alloc(cap_ * sizeof(float4));
        sorted_d_.alloc(cap_ * sizeof(float4));
        sort_hist_.alloc((size_t)kSlices * kSortBins * sizeof(int));
        sort_cur_.alloc((size_t)kSlices * kSortBins * sizeof(int));
        HIP_OK(hipMemset(sort_hist_.p, 0, sort_hist_.bytes));  // the scan clears it after every use
        HIP_OK(hipDeviceSynchronize());
    }
    ray_sort_ = on;
}

// Persistent grid of a stage kernel for a dynamic LDS size (cached).
template <bool POOLED>
Engine::Grids Engine::grids_for(int stack_entries) {
    const uint64_t key = ((uint64_t)stack_entries << 1) | (POOLED ? 1 : 0);
    auto it = grids_.find(key);
    if (it != grids_.end()) return it->second;
    const size_t lds = (size_t)stack_entries * kBlock * sizeof(uint32_t);
    Grids g;
    g.ext = persistent_grid(k_extend<false, POOLED, 0>, cus_, lds);
    g.ext_susp = persistent_grid(k_extend<false, POOLED, kSuspendLanes>, cus_, lds);
    g.ext_cont = persistent_grid(k_extend_cont<false, POOLED>, cus_, lds);
    g.shadow = persistent_grid(k_shadow<false, POOLED>, cus_, lds);
    g.shadow_susp = persistent_grid(k_shadow<false, POOLED, kShadowClosest, kSuspendLanes>, cus_, lds);
    g.shadow_cont = persistent_grid(k_shadow_cont<false, POOLED>, cus_, lds);
    g.shadow_early = persistent_grid(k_shadow<false, POOLED, kShadowEarly>, cus_, lds);
    g.shadow_segment = persistent_grid(k_shadow<false, POOLED, kShadowSegment>, cus_, lds);
    g.shade = persistent_grid(k_shade<POOLED>, cus_);
    g.shade_stack = persistent_grid(k_shade<POOLED>, cus_, lds);
    grids_[key] = g;
    return g;
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (int k = 0; k < kRing; ++k)
        if (ring_ev_[k]) (void)hipEventDestroy(ring_ev_[k]);
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    if (host_tables_) (void)hipHostFree(host_tables_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::set_grid_scale(float scale) {
    if (!(scale > 0.f)) throw std::invalid_argument("grid scale must be positive");
    if (scale == grid_scale_) return;
    HIP_OK(hipStreamSynchronize(stream_));
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    grid_scale_ = scale;
}

// explicit instantiations used by run()
template Engine::Grids Engine::grids_for<false>(int);
template Engine::Grids Engine::grids_for<true>(int);

void Engine::read_stats(unsigned long long out[kNumStats]) {
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipMemcpy(out, stats_.p, sizeof(unsigned long long) * kNumStats, hipMemcpyDeviceToHost));
    prof.collect();
}

void Engine::reset_stats() {
    HIP_OK(hipStreamSynchronize(stream_));
    k_zero_u64<<<1, 64, 0, stream_>>>(stats_.as<unsigned long long>(), kNumStats);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream_));
    prof.reset();
}

void Engine::run(const std::vector<PassItem>& items, bool count) {
    const int n = (int)items.size();
    if (n == 0) return;
    if (n > kMaxPassJobs) throw std::invalid_argument("too many jobs in one pass");
    // Fill the next ring slot: wait until the pass that used it last has consumed it.
    const int slot = ring_next_;
    ring_next_ = (ring_next_ + 1) % kRing;
    HIP_OK(hipEventSynchronize(ring_ev_[slot]));
    char* host = static_cast<char*>(host_tables_) + kPassTableBytes * slot;
    PassHeader* hdr = reinterpret_cast<PassHeader*>(host);
    JobDev* jobs = reinterpret_cast<JobDev*>(host + kPassHeaderBytes);
    size_t paths = 0, pixels = 0;
    int max_bounces = 0, stack = 1;
    bool classes[kNumClasses] = {false, false, false, false, false};
    for (int k = 0; k < n; ++k) {
        const PassItem& it = items[k];
        if (it.spp < 1) throw std::invalid_argument("pass item without samples");
        JobDev& J = jobs[k];
        J.sc = it.scene->dev();
        J.sc.env_alias = it.env_alias && J.sc.env_alias_row ? 1 : 0;
        J.fp = it.fp;
        J.fp.n_paths = it.fp.wh * it.spp;
        J.fp.bounce = 0;
        J.path_base = (int)paths;
        J.pix_base = (int)pixels;
        J.flush_survivors = it.flush_survivors;
        J.pad = 0;
        J.acc_d = it.acc_d;
        J.acc_i = it.acc_i;
        paths += (size_t)J.fp.n_paths;
        pixels += (size_t)it.fp.wh;
        max_bounces = std::max(max_bounces, it.fp.max_bounces);
        stack = std::max(stack, it.scene->max_stack());
        for (int c = 0; c < kNumClasses; ++c) classes[c] = classes[c] || it.scene->has_class(c);
    }
    if (paths > cap_) throw std::invalid_argument("pass exceeds the engine's path capacity");
    if (pixels > (size_t)0x7fffffff) throw std::invalid_argument("pass has too many pixels");
    hdr->n_jobs = n;
    hdr->n_paths = (int)paths;
    hdr->n_pixels = (int)pixels;
    // Slices split this pass's paths evenly (not the engine's capacity), so a
    // pass smaller than the capacity still spreads over all 64 slices and XCDs.
    hdr->region = (int)((paths + kSlices - 1) / kSlices);
    // One device table: stream order puts this copy after the previous pass's
    // kernels, so every pass (and every replay of a captured pass) reads the
    // same address; the host copy stays untouched until the copy has run.
    HIP_OK(hipMemcpyAsync(tables_.p, host, kPassHeaderBytes + sizeof(JobDev) * n, hipMemcpyHostToDevice, stream_));
    HIP_OK(hipEventRecord(ring_ev_[slot], stream_));
    const bool pooled = n > 1;
    // Straggler suspension (k_extend / k_shadow) for large passes over deep trees.
    const bool suspend = stack >= kSuspendMinStack && paths >= kSuspendMinPaths;
    const bool suspend_shadow = stack >= kShadowSuspendMinStack && paths >= kSuspendMinPaths;
    const bool graph = use_graphs_ && !prof.on && !count;
    uint64_t key = 0;
    if (graph) {
        int mask = 0;
        for (int c = 0; c < kNumClasses; ++c) mask |= classes[c] ? (1 << c) : 0;
        key = ((uint64_t)max_bounces << 24) | ((uint64_t)stack << 16) | ((uint64_t)mask << 8) | ((uint64_t)shadow_mode_ << 32) |
              (suspend_shadow ? 8 : 0) | (suspend ? 4 : 0) | (ray_sort_ ? 2 : 0) | (pooled ? 1 : 0);
        auto it = graphs_.find(key);
        if (it != graphs_.end()) {
            HIP_OK(hipGraphLaunch(it->second, stream_));
            return;
        }
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    }
    try {
        if (!pooled) run_t<false>(max_bounces, classes, count, stack, suspend, suspend_shadow);
        else run_t<true>(max_bounces, classes, count, stack, suspend, suspend_shadow);
    } catch (...) {
        if (graph) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        throw;
    }
    if (graph) {
        // The whole pass (raygen, bounce loop, resolve) becomes one graph,
        // replayed for every later pass of the same shape.
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_OK(e);
        graphs_[key] = exec;
        HIP_OK(hipGraphLaunch(exec, stream_));
    }
}

template <bool POOLED>
void Engine::run_t(int max_bounces, const bool* classes, bool count, int stack_entries, bool suspend,
                   bool suspend_shadow) {
    const char* dev = tables_.as<char>();
    const PassHeader* hdr = reinterpret_cast<const PassHeader*>(dev);
    const JobDev* jobs = reinterpret_cast<const JobDev*>(dev + kPassHeaderBytes);
    int* ctr = counters_.as<int>();
    unsigned long long* st = stats_.as<unsigned long long>();
    const int cap = (int)cap_;
    const Grids G = grids_for<POOLED>(stack_entries);
    const size_t lds = (size_t)stack_entries * kBlock * sizeof(uint32_t);
    auto scaled = [&](int g) { return std::max(1, (int)((float)g * grid_scale_ + 0.5f)); };

    const int gext = scaled(suspend ? G.ext_susp : G.ext);
    const int gext_cont = scaled(G.ext_cont);
    const int gsh_susp = scaled(G.shadow_susp);
    const int gsh_cont = scaled(G.shadow_cont);
    ContBufs cont;
    cont.hit = cont_hit_.as<float4>();
    cont.state = cont_state_.as<int4>();
    cont.stack = cont_stack_.as<uint32_t>();
    cont.room = cont_room_;
    cont.total = (size_t)cont_room_ * kSlices;
    const int gsh = scaled(shadow_mode_ == kShadowEarly     ? G.shadow_early
                           : shadow_mode_ == kShadowSegment ? G.shadow_segment
                                                            : G.shadow);
    // The shade launch carries the traversal stacks only when a subsurface class needs them.
    const int gshade = scaled(classes[kClsSubsurface] ? G.shade_stack : G.shade);
    const size_t shade_lds = classes[kClsSubsurface] ? lds : 0;
    int class_bits = 0;
    for (int c = 0; c < kNumClasses; ++c) class_bits |= classes[c] ? (1 << c) : 0;
    int q = 0;
    hipEvent_t b = prof.begin(stream_);
    k_raygen<POOLED><<<grid_linear_ * 4, kBlock, 0, stream_>>>(hdr, jobs, qo_[q].as<float4>(), qd_[q].as<float4>(),
                                                               qw_[q].as<float4>(), slot_d_.as<float4>(),
                                                               slot_i_.as<float4>(), ctr, q);
    HIP_OK(hipGetLastError());
    prof.end(kStageRaygen, b, stream_);
    for (int bounce = 0; bounce < max_bounces; ++bounce) {
        const int* perm = nullptr;
        if (ray_sort_ && bounce > 0) {  // primary rays are already in pixel order
            b = prof.begin(stream_);
            const dim3 g(std::max(1, grid_linear_ / kSlices), kSlices);
            k_sort_count<<<g, kBlock, 0, stream_>>>(keys_.as<uint16_t>(), hdr, ctr, q, sort_hist_.as<int>());
            k_sort_scan<<<kSlices, kSortBins, 0, stream_>>>(sort_hist_.as<int>(), sort_cur_.as<int>());
            k_sort_scatter<<<g, kBlock, 0, stream_>>>(keys_.as<uint16_t>(), hdr, ctr, q, sort_cur_.as<int>(),
                                                      perm_.as<int>());
            HIP_OK(hipGetLastError());
            prof.end(kStageSort, b, stream_);
            perm = perm_.as<int>();
        }
        b = prof.begin(stream_);
        auto extend = [&](auto count_tag) {
            constexpr bool C = decltype(count_tag)::value;
            float4* so = perm ? sorted_o_.as<float4>() : nullptr;
            float4* sd = perm ? sorted_d_.as<float4>() : nullptr;
            auto ext = suspend ? k_extend<C, POOLED, kSuspendLanes> : k_extend<C, POOLED, 0>;
            ext<<<gext, kBlock, lds, stream_>>>(jobs, qo_[q].as<float4>(), qd_[q].as<float4>(), qw_[q].as<float4>(),
                                                hits_.as<float4>(), lists_.as<int>(), cap, slot_d_.as<float4>(),
                                                slot_i_.as<float4>(), hdr, ctr, q, st, perm, so, sd, cont);
            if (!suspend) return;
            k_extend_cont<C, POOLED><<<gext_cont, kBlock, lds, stream_>>>(
                jobs, qo_[q].as<float4>(), qd_[q].as<float4>(), qw_[q].as<float4>(), hits_.as<float4>(),
                lists_.as<int>(), cap, slot_d_.as<float4>(), slot_i_.as<float4>(), hdr, ctr, st, perm, so, sd, cont);
        };
        if (count) extend(std::true_type());
        else extend(std::false_type());
        HIP_OK(hipGetLastError());
        prof.end(kStageExtend, b, stream_);
        b = prof.begin(stream_);
        // Classes absent from every job of the pass are skipped; their lists stay empty.
        k_shade<POOLED><<<gshade, kBlock, shade_lds, stream_>>>(
            jobs, bounce, class_bits, qo_[q].as<float4>(), qd_[q].as<float4>(), qw_[q].as<float4>(),
            hits_.as<float4>(), lists_.as<int>(), cap_, qo_[q ^ 1].as<float4>(), qd_[q ^ 1].as<float4>(),
            qw_[q ^ 1].as<float4>(), shadow_.as<float4>(), slot_d_.as<float4>(), slot_i_.as<float4>(), hdr, ctr, q,
            ray_sort_ ? keys_.as<uint16_t>() : nullptr, perm, perm ? sorted_o_.as<float4>() : nullptr,
            perm ? sorted_d_.as<float4>() : nullptr, st);
        HIP_OK(hipGetLastError());
        prof.end(kStageShade, b, stream_);
        b = prof.begin(stream_);
        auto shadow = [&](auto count_tag, auto mode_tag) {
            constexpr bool C = decltype(count_tag)::value;
            constexpr int M = decltype(mode_tag)::value;
            if (M == kShadowClosest && suspend_shadow) {
                k_shadow<C, POOLED, kShadowClosest, kSuspendLanes><<<gsh_susp, kBlock, lds, stream_>>>(
                    jobs, shadow_.as<float4>(), slot_d_.as<float4>(), slot_i_.as<float4>(), hdr, ctr, st, cont);
                k_shadow_cont<C, POOLED><<<gsh_cont, kBlock, lds, stream_>>>(
                    jobs, shadow_.as<float4>(), slot_d_.as<float4>(), slot_i_.as<float4>(), ctr, st, cont);
            } else {
                k_shadow<C, POOLED, M><<<gsh, kBlock, lds, stream_>>>(
                    jobs, shadow_.as<float4>(), slot_d_.as<float4>(), slot_i_.as<float4>(), hdr, ctr, st, cont);
            }
        };
        auto shadow_mode = [&](auto count_tag) {
            if (shadow_mode_ == kShadowEarly) shadow(count_tag, std::integral_constant<int, kShadowEarly>());
            else if (shadow_mode_ == kShadowSegment) shadow(count_tag, std::integral_constant<int, kShadowSegment>());
            else shadow(count_tag, std::integral_constant<int, kShadowClosest>());
        };
        if (count) shadow_mode(std::true_type());
        else shadow_mode(std::false_type());
        HIP_OK(hipGetLastError());
        prof.end(kStageShadow, b, stream_);
        q ^= 1;
    }
    b = prof.begin(stream_);
    k_resolve<POOLED><<<grid_linear_ * 4, kBlock, 0, stream_>>>(hdr, jobs, slot_d_.as<float4>(), slot_i_.as<float4>());
    HIP_OK(hipGetLastError());
    prof.end(kStageResolve, b, stream_);
}

FrameParams frame_params(const CameraRec& cam, const RenderSettings& set, uint32_t sample0, int spp) {
    FrameParams fp{};
    fp.cam_pos = cam.position;
    fp.left_down = cam.left_down;
    fp.du = cam.du;
    fp.dv = cam.dv;
    fp.width = cam.width;
    fp.height = cam.height;
    fp.wh = cam.width * cam.height;
    fp.seed = set.seed;
    fp.rr = set.rr;
    fp.rr_start = set.rr_start_depth;
    fp.max_bounces = set.max_bounces;
    fp.sample0 = sample0;
    fp.n_paths = fp.wh * spp;
    fp.bounce = 0;
    return fp;
}

// ------------------------------------------------------------------ frame state
size_t frame_buffer_bytes(int which, int width, int height) {
    size_t wh = (size_t)width * height;
    switch (which) {
        case kFrameMomentsDirect:
        case kFrameMomentsIndirect:
        case kFrameGNrmMat:
        case kFrameGBaryMesh: return wh * 8;
        case kFrameHistoryLength: return wh * 4;
        default: return (which >= 0 && which < kNumFrameBuffers) ? wh * 16 : 0;
    }
}

FrameState::FrameState(int width, int height) : w_(width), h_(height), wh_(width * height) {
    size_t wh = (size_t)wh_;
    for (int k = 0; k < 2; ++k) {
        g_pos_[k].alloc(wh * 16);
        g_nrm_[k].alloc(wh * 8);
        g_bary_[k].alloc(wh * 8);
        g_md_[k].alloc(wh * 16);
        h_d_[k].alloc(wh * 16);
        h_i_[k].alloc(wh * 16);
        m_d_[k].alloc(wh * 8);
        m_i_[k].alloc(wh * 8);
        h_len_[k].alloc(wh * 4);
        f_d_[k].alloc(wh * 16);
        f_i_[k].alloc(wh * 16);
    }
    v_d_.alloc(wh * 16);
    v_i_.alloc(wh * 16);
    denoised_.alloc(wh * 16);
    guide_nd_.alloc(wh * 16);
    guide_kl_.alloc(wh * 8);
    lum_.alloc(wh * 16);
    // One list of ceil(tiles / 8) entries per XCD band, for the tile shape
    // (set_tile_shift: 3..7) with the most tiles.
    size_t max_tiles = 0;
    for (int ts = 3; ts <= 7; ++ts) {
        const size_t tw = (size_t)1 << ts, th = (size_t)(kBlock >> ts);
        max_tiles = std::max(max_tiles, ((size_t)w_ + tw - 1) / tw * (((size_t)h_ + th - 1) / th));
    }
    short_tiles_.alloc((size_t)banded_grid((int)std::max<size_t>(max_tiles, 1)) * sizeof(int));
    short_n_.alloc(2 * kXcds * sizeof(int));
    HIP_OK(hipMemset(short_n_.p, 0, 2 * kXcds * sizeof(int)));
    HIP_OK(hipEventCreateWithFlags(&ev_order_, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_gbuf_, hipEventDisableTiming));
}

FrameState::~FrameState() {
    if (ev_order_) (void)hipEventDestroy(ev_order_);
    if (ev_gbuf_) (void)hipEventDestroy(ev_gbuf_);
}

size_t FrameState::device_bytes() const { return (size_t)wh_ * kFrameStateBytesPerPixel; }

static GBufferView gview(DevBuf* pos, DevBuf* nrm, DevBuf* bary, DevBuf* md, int k) {
    GBufferView v;
    v.pos_tri = pos[k].as<float4>();
    v.nrm_mat = nrm[k].as<ushort4>();
    v.bary_mesh = bary[k].as<ushort4>();
    v.motion_depth = md[k].as<float4>();
    return v;
}

// Matrices and planes of the G-buffer pass (reference: renderer.cuh:1463-1474):
// near 0.1, far = scene diagonal + distance from the camera to the scene centre,
// the previous view with the current projection.
static GBufferParams gbuffer_params(const HostScene& h, const CameraRec& cur, const CameraRec& prev) {
    GBufferParams g;
    g.cam_pos = cur.position;
    g.left_down = cur.left_down;
    g.du = cur.du;
    g.dv = cur.dv;
    const Box& b = h.bvh.bounds;
    vec3 center = (b.lo + b.hi) * 0.5f;
    g.near_plane = 0.1f;
    g.far_plane = length(b.hi - b.lo) + length(cur.position - center);
    camera_view_projection(cur, g.near_plane, g.far_plane, g.vp);
    camera_view_projection(prev, g.near_plane, g.far_plane, g.prev_vp);
    g.width = cur.width;
    g.height = cur.height;
    return g;
}

void FrameState::begin(const GpuScene& gs, const HostScene& host, const CameraRec& cam, float4* acc_d,
                       float4* acc_i, Profiler& prof, hipStream_t s, hipStream_t gbuf_stream) {
    if (cam.width != w_ || cam.height != h_) throw std::invalid_argument("camera resolution differs from the view's");
    const int c = fcur_;
    const bool first = frames_ == 0;
    const int ts = tile_shift_, tw = 1 << ts, th = kBlock >> ts;
    const int tiles = ((w_ + tw - 1) / tw) * ((h_ + th - 1) / th);
    // Slot c was the previous frame's "previous G-buffer": a G-buffer stream
    // first waits for everything enqueued on `s` so far.
    hipStream_t gsm = gbuf_stream ? gbuf_stream : s;
    if (gbuf_stream) {
        HIP_OK(hipEventRecord(ev_order_, s));
        HIP_OK(hipStreamWaitEvent(gsm, ev_order_, 0));
    }
    hipEvent_t e = prof.begin(gsm);
    GBufferParams g = gbuffer_params(host, cam, first ? cam : last_cam_);
    GBufferView v = gview(g_pos_, g_nrm_, g_bary_, g_md_, c);
    const size_t lds = (size_t)gs.max_stack() * kBlock * sizeof(uint32_t);
    k_gbuffer<<<tiles, kBlock, lds, gsm>>>(gs.dev(), g, v, ts);
    HIP_OK(hipGetLastError());
    prof.end(kStageGBuffer, e, gsm);
    if (gbuf_stream) {
        HIP_OK(hipEventRecord(ev_gbuf_, gsm));
        gbuf_pending_ = true;
    }
    // Fresh samples of this frame only (the reference resets its frame buffer
    // every denoised frame: renderer.cuh:1297-1302, 1121-1124).
    int grid = std::max(1, std::min((wh_ + kBlock - 1) / kBlock, 2048));
    k_zero4<<<grid, kBlock, 0, s>>>(acc_d, (size_t)wh_);
    k_zero4<<<grid, kBlock, 0, s>>>(acc_i, (size_t)wh_);
    HIP_OK(hipGetLastError());
    last_cam_ = cam;
    started_ = true;
}

void FrameState::end(int spp, const SvgfParams& p, const float4* acc_d, const float4* acc_i, Profiler& prof,
                     hipStream_t s) {
    if (!started_) throw std::logic_error("frame end without begin");
    started_ = false;
    if (gbuf_pending_) {
        HIP_OK(hipStreamWaitEvent(s, ev_gbuf_, 0));
        gbuf_pending_ = false;
    }
    const int c = fcur_, pv = c ^ 1;
    const bool first = frames_ == 0;
    const int ts = tile_shift_, tw = 1 << ts, th = kBlock >> ts;
    const int tiles = ((w_ + tw - 1) / tw) * ((h_ + th - 1) / th);
    GBufferView gc = gview(g_pos_, g_nrm_, g_bary_, g_md_, c), gp = gview(g_pos_, g_nrm_, g_bary_, g_md_, pv);
    hipEvent_t e = prof.begin(s);
    TemporalIO tio;
    tio.acc_d = acc_d;
    tio.acc_i = acc_i;
    tio.inv_spp = 1.f / (float)spp;
    tio.cur_d = h_d_[c].as<float4>();
    tio.cur_i = h_i_[c].as<float4>();
    tio.var_d = v_d_.as<float4>();
    tio.var_i = v_i_.as<float4>();
    tio.hist_d = h_d_[pv].as<float4>();
    tio.hist_i = h_i_[pv].as<float4>();
    tio.mom_prev_d = m_d_[pv].as<float2>();
    tio.mom_prev_i = m_i_[pv].as<float2>();
    tio.len_prev = h_len_[pv].as<int>();
    tio.mom_d = m_d_[c].as<float2>();
    tio.mom_i = m_i_[c].as<float2>();
    tio.len = h_len_[c].as<int>();
    tio.guide_nd = guide_nd_.as<float4>();
    tio.guide_kl = guide_kl_.as<int2>();
    tio.lum = lum_.as<float4>();
    int* n_short = short_n_.as<int>();
    const int ls = short_sel_;
    k_temporal<<<banded_grid(tiles), kBlock, 0, s>>>(w_, h_, ts, first ? 1 : 0, p, tio, gc, gp,
                                                     short_tiles_.as<int>(), n_short + kXcds * ls);
    HIP_OK(hipGetLastError());
    prof.end(kStageTemporal, e, s);

    e = prof.begin(s);
    FilterIO fio;
    fio.in_d = h_d_[c].as<float4>();
    fio.in_i = h_i_[c].as<float4>();
    fio.out_d = v_d_.as<float4>();
    fio.out_i = v_i_.as<float4>();
    fio.mom_d = m_d_[c].as<float2>();
    fio.mom_i = m_i_[c].as<float2>();
    fio.len = h_len_[c].as<int>();
    fio.guide_nd = guide_nd_.as<float4>();
    fio.guide_kl = guide_kl_.as<int2>();
    fio.lum = lum_.as<float4>();
    k_variance<<<banded_grid(tiles), kBlock, 0, s>>>(w_, h_, ts, p, fio, gc, short_tiles_.as<int>(),
                                                     n_short + kXcds * ls, n_short + kXcds * (ls ^ 1));
    HIP_OK(hipGetLastError());
    prof.end(kStageVariance, e, s);
    short_sel_ = ls ^ 1;

    out_sel_ = -1;
    if (!first) {
        // a-trous passes with step 1, 2, 4, ... (reference: renderer.cuh:1236-1250).
        int passes = std::max(1, std::min(8, p.atrous_iters));
        e = prof.begin(s);
        for (int k = 0; k < passes; ++k) {
            int o = k & 1;
            fio.in_d = k == 0 ? v_d_.as<float4>() : f_d_[o ^ 1].as<float4>();
            fio.in_i = k == 0 ? v_i_.as<float4>() : f_i_[o ^ 1].as<float4>();
            fio.out_d = f_d_[o].as<float4>();
            fio.out_i = f_i_[o].as<float4>();
            if (k == 0) {
                const int tx = (w_ + kAtTile - 1) / kAtTile, ty = (h_ + kAtTile - 1) / kAtTile;
                k_atrous_tile1<<<banded_grid(tx * ty), kBlock, 0, s>>>(w_, h_, tx, ty, p, fio, gc);
            } else {
                k_atrous<<<banded_grid(tiles), kBlock, 0, s>>>(w_, h_, ts, 1 << k, p, fio, gc);
            }
            HIP_OK(hipGetLastError());
            out_sel_ = o;
        }
        prof.end(kStageAtrous, e, s);
    }
    const float4* od = out_sel_ < 0 ? v_d_.as<float4>() : f_d_[out_sel_].as<float4>();
    const float4* oi = out_sel_ < 0 ? v_i_.as<float4>() : f_i_[out_sel_].as<float4>();
    int grid = std::max(1, std::min((wh_ + kBlock - 1) / kBlock, 2048));
    k_combine<<<grid, kBlock, 0, s>>>(od, oi, wh_, denoised_.as<float4>());
    HIP_OK(hipGetLastError());
    flast_ = c;
    fcur_ = pv;
    frames_++;
}

void FrameState::read(int which, void* out, hipStream_t s) {
    if (which < 0 || which >= kNumFrameBuffers) throw std::invalid_argument("unknown frame buffer");
    if (flast_ < 0) throw std::invalid_argument("no frame has been rendered");
    HIP_OK(hipStreamSynchronize(s));
    const int c = flast_;
    const DevBuf* src = nullptr;
    switch (which) {
        case kFrameDenoised: src = &denoised_; break;
        case kFrameDirect: src = out_sel_ < 0 ? &v_d_ : &f_d_[out_sel_]; break;
        case kFrameIndirect: src = out_sel_ < 0 ? &v_i_ : &f_i_[out_sel_]; break;
        case kFrameHistoryDirect: src = &h_d_[c]; break;
        case kFrameHistoryIndirect: src = &h_i_[c]; break;
        case kFrameMomentsDirect: src = &m_d_[c]; break;
        case kFrameMomentsIndirect: src = &m_i_[c]; break;
        case kFrameHistoryLength: src = &h_len_[c]; break;
        case kFrameVarianceDirect: src = &v_d_; break;
        case kFrameVarianceIndirect: src = &v_i_; break;
        case kFrameGPosTri: src = &g_pos_[c]; break;
        case kFrameGNrmMat: src = &g_nrm_[c]; break;
        case kFrameGBaryMesh: src = &g_bary_[c]; break;
        case kFrameGMotionDepth: src = &g_md_[c]; break;
    }
    HIP_OK(hipMemcpy(out, src->p, frame_buffer_bytes(which, w_, h_), hipMemcpyDeviceToHost));
}

// ------------------------------------------------------------------ renderer
struct Renderer::Impl {
    const GpuScene* scene = nullptr;
    const HostScene* host = nullptr;
    RenderSettings set;
    int width = 0, height = 0, wh = 0;
    int batch = 1;
    std::unique_ptr<Engine> eng;
    DevBuf acc_d, acc_i;
    int spp_done = 0;
    uint32_t sample_base = 0;  // sample index of the first sample in the accumulators
    CameraRec cam;             // camera of the following passes
    std::unique_ptr<FrameState> frames;
    hipStream_t gbuf_stream = nullptr;  // real-time frames: G-buffer beside the path tracing
};

Renderer::Renderer(const GpuScene& scene, const HostScene& host, const RenderSettings& s) : impl_(new Impl) {
    Impl& I = *impl_;
    I.scene = &scene;
    I.host = &host;
    I.set = s;
    I.width = host.camera.width;
    I.height = host.camera.height;
    I.wh = I.width * I.height;
    I.cam = host.camera;
    // Paths per pass: the scene's samples per pixel (its JSON "spp", at most 64)
    // of every pixel when they fit the queue budget, so a renderer reserves
    // queues for what its scene asks for (a 1-spp real-time renderer at 1080p:
    // 0.5 GB).  16 GiB by default: a 1920x1080 job of up to 32 samples per pixel
    // runs as one pass.  Renders of more samples run in several passes.
    size_t budget = s.queue_budget_bytes ? s.queue_budget_bytes : ((size_t)16 << 30);
    const size_t per_path = kQueueBytesPerPath + (s.ray_sort ? kSortBytesPerPath : 0);
    int bmax = (int)std::max<size_t>(1, budget / (per_path * (size_t)I.wh));
    int want = s.batch_spp > 0 ? s.batch_spp : std::min(64, std::max(1, host.spp));
    I.batch = std::max(1, std::min(want, bmax));
    if ((size_t)I.wh * I.batch > (size_t)0x7fffffff) I.batch = std::max(1, (int)(0x7fffffff / I.wh));
    I.eng.reset(new Engine((size_t)I.wh * I.batch, s.profile != 0));
    I.eng->set_use_graphs(s.use_graphs != 0);
    I.eng->set_ray_sort(s.ray_sort != 0);
    I.eng->set_shadow_mode(s.shadow_any_hit);
    I.acc_d.alloc((size_t)I.wh * 16);
    I.acc_i.alloc((size_t)I.wh * 16);
    HIP_OK(hipMemset(I.acc_d.p, 0, (size_t)I.wh * 16));
    HIP_OK(hipMemset(I.acc_i.p, 0, (size_t)I.wh * 16));
    HIP_OK(hipDeviceSynchronize());
}

Renderer::~Renderer() {
    if (impl_) {
        if (impl_->eng) (void)hipStreamSynchronize(impl_->eng->stream());
        if (impl_->gbuf_stream) {
            (void)hipStreamSynchronize(impl_->gbuf_stream);
            (void)hipStreamDestroy(impl_->gbuf_stream);
        }
        delete impl_;
    }
}

int Renderer::width() const { return impl_->width; }
int Renderer::height() const { return impl_->height; }
int Renderer::batch() const { return impl_->batch; }
int Renderer::spp_done() const { return impl_->spp_done; }

void Renderer::render(int spp, bool count) {
    Impl& I = *impl_;
    int remaining = spp;
    while (remaining > 0) {
        int k = std::min(remaining, I.batch);
        PassItem it;
        it.scene = I.scene;
        it.fp = frame_params(I.cam, I.set, I.sample_base + (uint32_t)I.spp_done, k);
        it.spp = k;
        it.flush_survivors = I.set.flush_survivors;
        it.env_alias = I.set.env_alias;
        it.acc_d = I.acc_d.as<float4>();
        it.acc_i = I.acc_i.as<float4>();
        I.eng->run({it}, count);
        I.spp_done += k;
        remaining -= k;
    }
}

void Renderer::synchronize() { HIP_OK(hipStreamSynchronize(impl_->eng->stream())); }

void Renderer::reset() {
    Impl& I = *impl_;
    hipStream_t s = I.eng->stream();
    int grid = std::max(1, std::min((I.wh + kBlock - 1) / kBlock, 2048));
    k_zero4<<<grid, kBlock, 0, s>>>(I.acc_d.as<float4>(), (size_t)I.wh);
    k_zero4<<<grid, kBlock, 0, s>>>(I.acc_i.as<float4>(), (size_t)I.wh);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    I.spp_done = 0;
    I.sample_base = 0;
}

void Renderer::reset_stats() { impl_->eng->reset_stats(); }

void Renderer::stats(RenderStats& out) {
    Impl& I = *impl_;
    unsigned long long s[kNumStats];
    I.eng->read_stats(s);
    out.ext_rays = s[kStatExtRays];
    out.shadow_rays = s[kStatShadowRays];
    out.ext_nodes = s[kStatExtNodes];
    out.ext_tris = s[kStatExtTris];
    out.shadow_nodes = s[kStatShadowNodes];
    out.shadow_tris = s[kStatShadowTris];
    out.probe_rays = s[kStatProbeRays];
    out.miss_paths = s[kStatMisses];
    out.shadow_lit = s[kStatShadowLit];
    for (int k = 0; k < kNumClasses; ++k) out.class_paths[k] = s[kStatClassPaths + k];
    for (int k = 0; k < kNumStages; ++k) {
        out.stage_ms[k] = I.eng->prof.ms[k];
        out.stage_launches[k] = I.eng->prof.launches[k];
    }
    out.spp_done = I.spp_done;
    out.batch_spp = I.batch;
}

void read_accum_buffers(const float4* acc_d, const float4* acc_i, int wh, int spp_done, int which, float* out,
                        hipStream_t s) {
    HIP_OK(hipStreamSynchronize(s));
    size_t n = (size_t)wh * 4;
    if (which == 0) {
        HIP_OK(hipMemcpy(out, acc_d, n * 4, hipMemcpyDeviceToHost));
        return;
    }
    if (which == 1) {
        HIP_OK(hipMemcpy(out, acc_i, n * 4, hipMemcpyDeviceToHost));
        return;
    }
    std::vector<float> d(n), ind(n);
    HIP_OK(hipMemcpy(d.data(), acc_d, n * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ind.data(), acc_i, n * 4, hipMemcpyDeviceToHost));
    int spp = spp_done > 0 ? spp_done : 1;
    for (size_t k = 0; k < (size_t)wh; ++k) {
        for (int c = 0; c < 3; ++c) out[4 * k + c] = (d[4 * k + c] + ind[4 * k + c]) / (float)spp;
        out[4 * k + 3] = 1.f;
    }
}

void Renderer::read_accum(int which, float* out) {
    Impl& I = *impl_;
    read_accum_buffers(I.acc_d.as<float4>(), I.acc_i.as<float4>(), I.wh, I.spp_done, which, out, I.eng->stream());
}

void Renderer::set_camera(const CameraRec& c) {
    Impl& I = *impl_;
    if (c.width != I.width || c.height != I.height) throw std::invalid_argument("camera resolution differs from the renderer's");
    I.cam = c;
    reset();
}

const CameraRec& Renderer::camera() const { return impl_->cam; }
int Renderer::frames() const { return impl_->frames ? impl_->frames->frames() : 0; }

void Renderer::reset_history() {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.eng->stream()));
    if (I.frames) I.frames->reset_history();
}

void Renderer::render_frame(int spp, const SvgfParams& p) {
    Impl& I = *impl_;
    if (spp < 1) throw std::invalid_argument("spp must be positive");
    if (!I.frames) I.frames.reset(new FrameState(I.width, I.height));
    if (!I.gbuf_stream) HIP_OK(hipStreamCreateWithFlags(&I.gbuf_stream, hipStreamNonBlocking));
    hipStream_t s = I.eng->stream();
    I.frames->begin(*I.scene, *I.host, I.cam, I.acc_d.as<float4>(), I.acc_i.as<float4>(), I.eng->prof, s,
                    I.gbuf_stream);
    I.spp_done = 0;
    I.sample_base = I.frames->next_sample0(spp);
    render(spp);
    I.frames->end(spp, p, I.acc_d.as<float4>(), I.acc_i.as<float4>(), I.eng->prof, s);
}

void Renderer::read_frame(int which, void* out) {
    Impl& I = *impl_;
    if (!I.frames) throw std::invalid_argument("no frame has been rendered");
    I.frames->read(which, out, I.eng->stream());
}

// ------------------------------------------------------------------ tools
void trace_batch(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts) {
    DevBuf r, o, c;
    r.alloc((size_t)n * 32);
    o.alloc((size_t)n * 16);
    c.alloc((size_t)n * 8);
    HIP_OK(hipMemcpy(r.p, rays, (size_t)n * 32, hipMemcpyHostToDevice));
    int grid = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
    const size_t lds = (size_t)scene.max_stack() * kBlock * sizeof(uint32_t);
    k_trace_batch<<<grid, kBlock, lds>>>(scene.dev(), r.as<float4>(), n, o.as<float4>(), c.as<uint2>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out_hits, o.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (out_counts) HIP_OK(hipMemcpy(out_counts, c.p, (size_t)n * 8, hipMemcpyDeviceToHost));
}

// Staging of trace_batch_noalloc: the rays (32 B each), then the hits (16 B)
// and the traversal counts (8 B) of up to kNoAllocMaxRays rays.  Each GPU the
// library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[(size_t)kNoAllocMaxRays * 56];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device
// (the array's name is only a host-side handle; kernels and copies need this).
void* scratch_on_current_device() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("HIP error: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void trace_batch_noalloc(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts) {
    if (n < 0 || n > kNoAllocMaxRays)
        throw std::invalid_argument("rt_trace_rays_noalloc takes at most " + std::to_string(kNoAllocMaxRays) +
                                    " rays per call, got " + std::to_string(n));
    if (n == 0) return;
    unsigned char* base = (unsigned char*)scratch_on_current_device();
    float4* r = (float4*)base;                               // 2 float4 per ray
    float4* o = (float4*)(base + (size_t)n * 32);            // hits
    uint2* c = (uint2*)(base + (size_t)n * 48);              // counts
    HIP_OK(hipMemcpy(r, rays, (size_t)n * 32, hipMemcpyHostToDevice));
    int grid = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
    const size_t lds = (size_t)scene.max_stack() * kBlock * sizeof(uint32_t);
    k_trace_batch<<<grid, kBlock, lds>>>(scene.dev(), r, n, o, c);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out_hits, o, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (out_counts) HIP_OK(hipMemcpy(out_counts, c, (size_t)n * 8, hipMemcpyDeviceToHost));
}

void texture_lookup(const GpuScene& scene, const HostScene& host, int tex, const float* uv, int n, float* rgb_out) {
    if (tex < 0 || tex >= (int)host.textures.size() || host.textures[tex].w <= 0)
        throw std::invalid_argument("no decoded texture at that index");
    DevBuf u, o;
    u.alloc((size_t)n * 8);
    o.alloc((size_t)n * 16);
    HIP_OK(hipMemcpy(u.p, uv, (size_t)n * 8, hipMemcpyHostToDevice));
    k_texture_lookup<<<(n + kBlock - 1) / kBlock, kBlock>>>(scene.dev(), tex, u.as<float2>(), n, o.as<float4>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    std::vector<float4> res((size_t)n);
    HIP_OK(hipMemcpy(res.data(), o.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        rgb_out[3 * i] = res[i].x;
        rgb_out[3 * i + 1] = res[i].y;
        rgb_out[3 * i + 2] = res[i].z;
    }
}

void math_selftest(int op, const float* a, const float* b, float* out, int n) {
    DevBuf da, db, dout;
    da.alloc((size_t)n * 4);
    db.alloc((size_t)n * 4);
    dout.alloc((size_t)n * 4);
    HIP_OK(hipMemcpy(da.p, a, (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(db.p, b, (size_t)n * 4, hipMemcpyHostToDevice));
    k_math<<<(n + 255) / 256, 256>>>(op, da.as<float>(), db.as<float>(), dout.as<float>(), n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost));
}

}  // namespace rt
