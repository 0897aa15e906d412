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
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "integrator.h"
#include "traverse.h"

namespace rt {

#define HIP_OK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                     #x);                                                           \
    } while (0)

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
struct SliceCursor {
    int* counters;
    int cnt_kind, fetch_kind;
    int slice, visited, count;
    __device__ void init(int* c, int ck, int fk) {
        counters = c;
        cnt_kind = ck;
        fetch_kind = fk;
        int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
        slice = wave % kSlices;
        visited = 0;
        count = *ctr_at(counters, cnt_kind, slice);
    }
    __device__ bool next(int& base) {
        while (visited < kSlices) {
            int* f = ctr_at(counters, fetch_kind, slice);
            int b = count;
            // Drained slices are skipped with a plain load; only live ones take an atomic.
            // Lane 0 decides and broadcasts, so the wave stays uniform.
            if ((threadIdx.x & (kWave - 1)) == 0 && load_relaxed(f) < count) b = atomicAdd(f, kWave);
            b = __shfl(b, 0);
            if (b < count) {
                base = b;
                return true;
            }
            ++visited;
            slice = slice + 1 == kSlices ? 0 : slice + 1;
            count = *ctr_at(counters, cnt_kind, slice);
        }
        return false;
    }
};

// ------------------------------------------------------------------ raygen
// Jittered pinhole ray per (pixel, sample) (reference: Src/Core/camera.cu:27-35,
// Src/Core/renderer.cuh:96-115).  Path p covers pixel p % wh of sample
// sample0 + p / wh; row 0 of the image is the bottom row.  Path p sits at
// queue index p, in slice p / region.
__global__ __launch_bounds__(kBlock) void k_raygen(FrameParams fp, float4* qo, float4* qd, float4* qw,
                                                    float4* slot_d, float4* slot_i, int* counters, int qslot,
                                                    int region) {
    int stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        int n = fp.n_paths - sl * region;
        *ctr_at(counters, qslot, sl) = n < 0 ? 0 : (n > region ? region : n);
        *ctr_at(counters, kFetchExtend, sl) = 0;
        for (int c = 0; c < kNumClasses; ++c) {
            *ctr_at(counters, kCntClass0 + c, sl) = 0;
            *ctr_at(counters, kFetchClass0 + c, sl) = 0;
        }
    }
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < fp.n_paths; p += stride) {
        uint32_t pixel = (uint32_t)(p % fp.wh);
        uint32_t sample = fp.sample0 + (uint32_t)(p / fp.wh);
        int i = (int)(pixel % (uint32_t)fp.width), j = (int)(pixel / (uint32_t)fp.width);
        Rng rng;
        rng.init(fp.seed, pixel, sample, 0u, kStreamCamera);
        float x0 = rng.next();
        float x1 = rng.next();
        vec3 pp = (fp.left_down + ((float)i + x0) * fp.du) + ((float)j + x1) * fp.dv;
        vec3 d = normalize(pp - fp.cam_pos);
        qo[p] = make_float4(fp.cam_pos.x, fp.cam_pos.y, fp.cam_pos.z, 0.f);
        qd[p] = make_float4(d.x, d.y, d.z, i_as_f((int)kFlagFirst));
        qw[p] = make_float4(1.f, 1.f, 1.f, i_as_f(p));
        slot_d[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        slot_i[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ------------------------------------------------------------------ extend
// Closest hit for every path of the current queue, then the path's queue
// index is appended to its slice's list for the shading class of what it hit.
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_extend(DevScene sc, const float4* __restrict__ qo,
                                                    const float4* __restrict__ qd, float4* __restrict__ hits,
                                                    int* __restrict__ lists, int cap, int region, int* counters,
                                                    int qslot, unsigned long long* stats) {
    __shared__ uint32_t stk[STACK * kBlock];
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        *ctr_at(counters, qslot ^ 1, sl) = 0;
        *ctr_at(counters, kCntShadow, sl) = 0;
        *ctr_at(counters, kFetchShadow, sl) = 0;
        int n = *ctr_at(counters, qslot, sl);
        if (n > 0) atomicAdd(&stats[kStatExtRays], (unsigned long long)n);
    }
    uint32_t* col = stk + threadIdx.x;
    SliceCursor cur;
    cur.init(counters, qslot, kFetchExtend);
    int base;
    while (cur.next(base)) {
        int j = base + (int)(threadIdx.x & (kWave - 1));
        int i = cur.slice * region + j;
        TravCount c{0, 0};
        int cls = -1;
        if (j < cur.count) {
            float4 o = qo[i], d = qd[i];
            RayInv r = make_ray(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
            Hit h = trace_closest<STACK, COUNT>(sc, r, 0.f, 1e8f, col, kBlock, c);
            hits[i] = make_float4(h.t, h.b1, h.b2, i_as_f(h.prim));
            cls = shade_class_of(sc.tri_attr, h.prim);
        }
#pragma unroll
        for (int k = 0; k < kNumClasses; ++k) {
            if (__ballot(cls == k) == 0ull) continue;
            int pos = wave_append(cls == k, ctr_at(counters, kCntClass0 + k, cur.slice));
            if (cls == k) lists[(size_t)k * cap + (size_t)cur.slice * region + pos] = i;
        }
        if (COUNT) {
            wave_add_u64(&stats[kStatExtNodes], c.nodes);
            wave_add_u64(&stats[kStatExtTris], c.tris);
        }
    }
}

// ------------------------------------------------------------------ shade
// One kernel per shading class; each reads the queue indices binned by extend
// and appends survivors / shadow records to the slice they came from.
template <int CLS, int STACK>
__global__ __launch_bounds__(kBlock) void k_shade(DevScene sc, FrameParams fp, const float4* __restrict__ qo,
                                                   const float4* __restrict__ qd, const float4* __restrict__ qw,
                                                   const float4* __restrict__ hits, const int* __restrict__ list,
                                                   float4* __restrict__ no, float4* __restrict__ nd,
                                                   float4* __restrict__ nw, float4* __restrict__ shadow,
                                                   float4* slot_d, float4* slot_i, int region, int* counters,
                                                   int qslot) {
    __shared__ uint32_t stk[CLS == kClsSubsurface ? STACK * kBlock : 1];
    if (blockIdx.x == 0 && threadIdx.x < kSlices) *ctr_at(counters, kFetchExtend, threadIdx.x) = 0;
    uint32_t* col = stk + (CLS == kClsSubsurface ? threadIdx.x : 0);
    SliceCursor cur;
    cur.init(counters, kCntClass0 + CLS, kFetchClass0 + CLS);
    int base;
    while (cur.next(base)) {
        int j = base + (int)(threadIdx.x & (kWave - 1));
        ShadeOut so;
        so.alive = false;
        so.shadow = false;
        PathState ps;
        int pid = 0;
        if (j < cur.count) {
            int i = list[(size_t)cur.slice * region + j];
            float4 o = qo[i], d = qd[i], w = qw[i];
            ps.o = v3(o.x, o.y, o.z);
            ps.pdf = o.w;
            ps.d = v3(d.x, d.y, d.z);
            ps.flags = (uint32_t)f_as_i(d.w);
            ps.w = v3(w.x, w.y, w.z);
            pid = f_as_i(w.w);
            uint32_t pixel = (uint32_t)(pid % fp.wh);
            uint32_t sample = fp.sample0 + (uint32_t)(pid / fp.wh);
            so = shade_path<CLS, STACK>(sc, fp, ps, hits[i], pixel, sample, col, kBlock);
            if (so.has_emit) {
                float4* s = so.add_direct ? slot_d : slot_i;
                float4 a = s[pid];
                a.x += so.emit.x;
                a.y += so.emit.y;
                a.z += so.emit.z;
                s[pid] = a;
            }
        }
        if (CLS != kClsTerminal) {
            int qpos = wave_append(so.alive, ctr_at(counters, qslot ^ 1, cur.slice));
            if (so.alive) {
                size_t o = (size_t)cur.slice * region + qpos;
                no[o] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.pdf);
                nd[o] = make_float4(ps.d.x, ps.d.y, ps.d.z, i_as_f((int)ps.flags));
                nw[o] = make_float4(ps.w.x, ps.w.y, ps.w.z, i_as_f(pid));
            }
        }
        if (CLS == kClsDiffuse || CLS == kClsPbr || CLS == kClsSubsurface) {
            int spos = wave_append(so.shadow, ctr_at(counters, kCntShadow, cur.slice));
            if (so.shadow) {
                float4* r = shadow + 4 * ((size_t)cur.slice * region + spos);
                r[0] = make_float4(so.rec.lp.x, so.rec.lp.y, so.rec.lp.z, so.rec.sdist);
                r[1] = make_float4(so.rec.dir.x, so.rec.dir.y, so.rec.dir.z, i_as_f(pid));
                r[2] = make_float4(so.rec.hp.x, so.rec.hp.y, so.rec.hp.z, i_as_f(so.add_direct ? 0 : 1));
                r[3] = make_float4(so.rec.contrib.x, so.rec.contrib.y, so.rec.contrib.z, 0.f);
            }
        }
    }
}

// ------------------------------------------------------------------ shadow
// Closest hit from the light point towards the shading point with tMin 1e-6,
// then the reference's 0.1 distance tolerances (Src/Core/light.cu:5-27).
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_shadow(DevScene sc, const float4* __restrict__ shadow, float4* slot_d,
                                                    float4* slot_i, int region, int* counters,
                                                    unsigned long long* stats) {
    __shared__ uint32_t stk[STACK * kBlock];
    if (blockIdx.x == 0 && threadIdx.x < kSlices) {
        int sl = threadIdx.x;
        int n = *ctr_at(counters, kCntShadow, sl);
        if (n > 0) atomicAdd(&stats[kStatShadowRays], (unsigned long long)n);
        // The class lists of this bounce are consumed: reset them for the next extend.
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
        if (j < cur.count) {
            const float4* r = shadow + 4 * ((size_t)cur.slice * region + j);
            float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            ShadowRec rec;
            rec.lp = v3(r0.x, r0.y, r0.z);
            rec.sdist = r0.w;
            rec.dir = v3(r1.x, r1.y, r1.z);
            rec.hp = v3(r2.x, r2.y, r2.z);
            vec3 org = rec.lp + 0.01f * rec.dir;
            RayInv ray = make_ray(org, rec.dir);
            Hit h = trace_closest<STACK, COUNT>(sc, ray, 1e-6f, 1e8f, col, kBlock, c);
            bool hit = h.prim >= 0;
            vec3 tp = v3(0.f, 0.f, 0.f);
            if (hit) tp = position_at(sc.tri_attr, h.prim, h.b1, h.b2);
            if (shadow_visible(rec, hit, tp)) {
                int pid = f_as_i(r1.w);
                float4* s = f_as_i(r2.w) == 0 ? slot_d : slot_i;
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
        }
    }
}

// Paths still alive after the last bounce are dropped without contributing
// (reference: Src/Core/renderer.cuh:1166-1194 never accumulates them).
__global__ __launch_bounds__(kBlock) void k_lost(const float4* __restrict__ qw, float4* slot_d, float4* slot_i,
                                                  int* counters, int qslot, int region) {
    for (int sl = 0; sl < kSlices; ++sl) {
        const int count = *ctr_at(counters, qslot, sl);
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
            int pid = f_as_i(qw[(size_t)sl * region + j].w);
            slot_d[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
            slot_i[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// Per-pixel accumulation in sample order, so the sum is the same as adding
// one sample per frame (reference: Src/Core/renderer.cuh:117-122).
__global__ __launch_bounds__(kBlock) void k_resolve(const float4* __restrict__ slot_d,
                                                     const float4* __restrict__ slot_i, int wh, int k,
                                                     float4* acc_d, float4* acc_i) {
    for (int px = blockIdx.x * blockDim.x + threadIdx.x; px < wh; px += gridDim.x * blockDim.x) {
        float4 a = acc_d[px], b = acc_i[px];
        for (int s = 0; s < k; ++s) {
            float4 x = slot_d[(size_t)s * wh + px];
            float4 y = slot_i[(size_t)s * wh + px];
            a.x += x.x; a.y += x.y; a.z += x.z;
            b.x += y.x; b.y += y.y; b.z += y.z;
        }
        acc_d[px] = a;
        acc_i[px] = b;
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

// Closest hit for an arbitrary ray batch: rays are (o, tmin), (d, tmax).
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_trace_batch(DevScene sc, const float4* __restrict__ rays, int n,
                                                         float4* __restrict__ out, uint2* __restrict__ counts) {
    __shared__ uint32_t stk[STACK * kBlock];
    uint32_t* col = stk + threadIdx.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 o = rays[2 * (size_t)i], d = rays[2 * (size_t)i + 1];
        RayInv r = make_ray(v3(o.x, o.y, o.z), v3(d.x, d.y, d.z));
        TravCount c{0, 0};
        Hit h = trace_closest<STACK, true>(sc, r, o.w, d.w, col, kBlock, c);
        out[i] = make_float4(h.t, h.b1, h.b2, i_as_f(h.prim));
        counts[i] = make_uint2(c.nodes, c.tris);
    }
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
// Per-pixel frame kernels work on 16x16 tiles (four 16x4 waves), so the
// neighbourhood taps of the filters stay within a few cache lines per wave.
__device__ __forceinline__ bool tile_pixel(int w, int h, int& i, int& j) {
    int tiles_x = (w + 15) >> 4;
    int t = blockIdx.x;
    i = (t % tiles_x) * 16 + (threadIdx.x & 15);
    j = (t / tiles_x) * 16 + (threadIdx.x >> 4);
    return i < w && j < h;
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_gbuffer(DevScene sc, GBufferParams g, GBufferView out) {
    __shared__ uint32_t stk[STACK * kBlock];
    int i, j;
    if (!tile_pixel(g.width, g.height, i, j)) return;
    gbuffer_pixel<STACK>(sc, g, i, j, stk + threadIdx.x, kBlock, out);
}

__global__ __launch_bounds__(kBlock) void k_temporal(int w, int h, int first, SvgfParams p, TemporalIO io,
                                                      GBufferView cur, GBufferView prev) {
    int i, j;
    if (!tile_pixel(w, h, i, j)) return;
    temporal_pixel(i, j, w, h, first != 0, p, io, cur, prev);
}

__global__ __launch_bounds__(kBlock) void k_variance(int w, int h, SvgfParams p, FilterIO io, GBufferView g) {
    int i, j;
    if (!tile_pixel(w, h, i, j)) return;
    variance_pixel(i, j, w, h, p, io, g);
}

__global__ __launch_bounds__(kBlock) void k_atrous(int w, int h, int step, SvgfParams p, FilterIO io, GBufferView g) {
    int i, j;
    if (!tile_pixel(w, h, i, j)) return;
    atrous_pixel(i, j, w, h, step, p, io, g);
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

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void alloc(size_t b) {
        release();
        if (b == 0) return;
        HIP_OK(hipMalloc(&p, b));
        bytes = b;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p); }
    ~DevBuf() { release(); }
};

struct GpuScene::Impl {
    DevBuf nodes, tri_isect, tri_attr, materials, lights, light_tris, env_rgb, env_row, env_marg;
    DevScene dev{};
    int stack = 32;
    bool has_class[kNumClasses] = {true, false, false, false, false};
    int device = 0;
};

static void upload(DevBuf& b, const void* src, size_t bytes) {
    b.alloc(bytes);
    if (bytes) HIP_OK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

GpuScene::GpuScene(const HostScene& sc) : impl_(new Impl) {
    Impl& I = *impl_;
    HIP_OK(hipGetDevice(&I.device));
    const BvhBuild& bvh = sc.bvh;
    int n = sc.num_tris();
    if (bvh.max_stack > 64) throw std::runtime_error("BVH deeper than the 64-entry traversal stack");
    I.stack = bvh.max_stack <= 32 ? 32 : 64;
    // Packed nodes; a leaf-only tree gets one dummy line so the pointer is valid.
    std::vector<PackedNode> nodes = bvh.packed;
    if (nodes.empty()) nodes.resize(1);
    upload(I.nodes, nodes.data(), nodes.size() * sizeof(PackedNode));
    std::vector<float4> isect(3 * (size_t)n), attr(kAttrStride * (size_t)n);
    auto class_of = [&](int material) -> int {
        switch (sc.materials[material].type) {
            case kLight: return kClsTerminal;
            case kPbr: return kClsPbr;
            case kDielectric: return kClsDielectric;
            case kSubsurface: return kClsSubsurface;
            default: return kClsDiffuse;
        }
    };
    for (int k = 0; k < n; ++k) {
        int t = bvh.order[k];
        const float* p = &sc.pos[9 * (size_t)t];
        const float* nn = &sc.nrm[9 * (size_t)t];
        const float* uv = &sc.uv[6 * (size_t)t];
        const float* g = &sc.gnormal[3 * (size_t)t];
        vec3 v0 = v3(p[0], p[1], p[2]), v1 = v3(p[3], p[4], p[5]), v2 = v3(p[6], p[7], p[8]);
        vec3 e1 = v1 - v0, e2 = v2 - v0;
        isect[3 * (size_t)k + 0] = make_float4(v0.x, v0.y, v0.z, 0.f);
        isect[3 * (size_t)k + 1] = make_float4(e1.x, e1.y, e1.z, 0.f);
        isect[3 * (size_t)k + 2] = make_float4(e2.x, e2.y, e2.z, 0.f);
        float4* a = &attr[kAttrStride * (size_t)k];
        const MeshRec& mr = sc.meshes[sc.mesh[t]];
        int cls = class_of(mr.material);
        I.has_class[cls] = true;
        a[0] = make_float4(p[0], p[1], p[2], i_as_f(t - mr.start));
        a[1] = make_float4(p[3], p[4], p[5], g[0]);
        a[2] = make_float4(p[6], p[7], p[8], g[1]);
        a[3] = make_float4(nn[0], nn[1], nn[2], g[2]);
        a[4] = make_float4(nn[3], nn[4], nn[5], uv[0]);
        a[5] = make_float4(nn[6], nn[7], nn[8], uv[1]);
        a[6] = make_float4(uv[2], uv[3], uv[4], uv[5]);
        a[7] = make_float4(i_as_f(sc.mesh[t]), i_as_f(mr.material), i_as_f(mr.light), i_as_f(cls));
    }
    upload(I.tri_isect, isect.data(), isect.size() * sizeof(float4));
    upload(I.tri_attr, attr.data(), attr.size() * sizeof(float4));
    std::vector<MaterialRec> mats = sc.materials;
    if (mats.empty()) mats.resize(1);
    upload(I.materials, mats.data(), mats.size() * sizeof(MaterialRec));
    std::vector<LightRec> lights = sc.lights;
    std::vector<float4> ltris;
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
    upload(I.lights, lights.data(), lights.size() * sizeof(LightRec));
    upload(I.light_tris, ltris.data(), ltris.size() * sizeof(float4));

    DevScene& d = I.dev;
    d.nodes = I.nodes.as<PackedNode>();
    d.root_ref = bvh.root_ref;
    d.tri_isect = I.tri_isect.as<float4>();
    d.tri_attr = I.tri_attr.as<float4>();
    d.materials = I.materials.as<MaterialRec>();
    d.lights = I.lights.as<LightRec>();
    d.num_lights = (int)sc.lights.size();
    d.env_light = sc.env_light;
    d.light_tris = I.light_tris.as<float4>();
    d.bounds_lo = bvh.bounds.lo;
    d.bounds_hi = bvh.bounds.hi;
    d.env_rgb = nullptr;
    d.env_w = d.env_h = 0;
    d.env_row_cdf = d.env_marg_cdf = d.env_pdf = nullptr;
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
    }
}

GpuScene::~GpuScene() { delete impl_; }

const DevScene& GpuScene::dev() const { return impl_->dev; }
int GpuScene::stack() const { return impl_->stack; }
bool GpuScene::has_class(int c) const { return impl_->has_class[c]; }

// Row-conditional and marginal CDFs weighted by sin(theta) (reference:
// Src/Core/texture.cu:38-80).
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
    for (int i = 1; i < h; ++i) t.marg_cdf[i] += t.marg_cdf[i - 1];
    float total = h > 0 ? t.marg_cdf[h - 1] : 0.f;
    for (int i = 0; i < h; ++i) t.marg_cdf[i] = total > 0.0f ? t.marg_cdf[i] / total : 0.0f;
}

// ------------------------------------------------------------------ renderer
struct Renderer::Impl {
    const GpuScene* scene = nullptr;
    const HostScene* host = nullptr;
    RenderSettings set;
    hipStream_t stream = nullptr;
    int width = 0, height = 0, wh = 0;
    int batch = 1;
    size_t cap = 0;  // paths per pass
    int region = 0;  // queue entries per slice
    DevBuf qo[2], qd[2], qw[2], hits, lists, shadow, slot_d, slot_i, acc_d, acc_i, counters, stats;
    int grid_persist_ext = 0, grid_persist_shadow = 0, grid_linear = 0;
    int grid_shade[kNumClasses] = {0};
    int spp_done = 0;
    uint32_t sample_base = 0;  // sample index of the first sample in the accumulators
    CameraRec cam;             // camera of the following passes
    // Real-time frames: two G-buffers / histories (this frame, previous frame).
    bool frame_alloc = false;
    DevBuf g_pos[2], g_nrm[2], g_bary[2], g_md[2];
    DevBuf h_d[2], h_i[2], m_d[2], m_i[2], h_len[2];
    DevBuf v_d, v_i, f_d[2], f_i[2], denoised;
    CameraRec last_frame_cam;
    int fcur = 0;     // buffer set the next frame writes
    int flast = -1;   // buffer set of the last frame (-1: none yet)
    int frames = 0;   // frames since the last history reset
    int out_sel = -1; // -1: the variance pass output is final, else f_*[out_sel]
    // timing
    bool profile = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    double stage_ms[kNumStages] = {0};
    int stage_launches[kNumStages] = {0};
};

template <typename K>
static int persistent_grid(K kernel, int cus) {
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0));
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

Renderer::Renderer(const GpuScene& scene, const HostScene& host, const RenderSettings& s) : impl_(new Impl) {
    Impl& I = *impl_;
    I.scene = &scene;
    I.host = &host;
    I.set = s;
    I.width = host.camera.width;
    I.height = host.camera.height;
    I.wh = I.width * I.height;
    I.cam = host.camera;
    I.profile = s.profile != 0;
    HIP_OK(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    int cus = 0;
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // Paths per pass: all requested samples of every pixel when they fit the
    // queue budget (about 200 bytes of queue state per path).
    size_t per_path = 3 * 16 * 2 + 16 + 4 * kNumClasses + 64 + 32;
    size_t budget = s.queue_budget_bytes ? s.queue_budget_bytes : ((size_t)8 << 30);
    int bmax = (int)std::max<size_t>(1, budget / (per_path * (size_t)I.wh));
    int want = s.batch_spp > 0 ? s.batch_spp : 64;
    I.batch = std::max(1, std::min(want, bmax));
    if ((size_t)I.wh * I.batch > (size_t)0x7fffffff) I.batch = std::max(1, (int)(0x7fffffff / I.wh));
    I.cap = (size_t)I.wh * I.batch;
    for (int k = 0; k < 2; ++k) {
        I.qo[k].alloc(I.cap * 16);
        I.qd[k].alloc(I.cap * 16);
        I.qw[k].alloc(I.cap * 16);
    }
    I.hits.alloc(I.cap * 16);
    I.lists.alloc(I.cap * 4 * kNumClasses);
    I.shadow.alloc(I.cap * 64);
    I.slot_d.alloc(I.cap * 16);
    I.slot_i.alloc(I.cap * 16);
    I.acc_d.alloc((size_t)I.wh * 16);
    I.acc_i.alloc((size_t)I.wh * 16);
    I.region = (int)((I.cap + kSlices - 1) / kSlices);
    I.counters.alloc(kCounterInts * sizeof(int));
    I.stats.alloc(kNumStats * sizeof(unsigned long long));
    HIP_OK(hipMemset(I.counters.p, 0, kCounterInts * sizeof(int)));
    HIP_OK(hipMemset(I.stats.p, 0, kNumStats * sizeof(unsigned long long)));
    HIP_OK(hipMemset(I.acc_d.p, 0, (size_t)I.wh * 16));
    HIP_OK(hipMemset(I.acc_i.p, 0, (size_t)I.wh * 16));
    if (scene.stack() == 32) {
        I.grid_persist_ext = persistent_grid(k_extend<32, false>, cus);
        I.grid_persist_shadow = persistent_grid(k_shadow<32, false>, cus);
        I.grid_shade[kClsSubsurface] = persistent_grid(k_shade<kClsSubsurface, 32>, cus);
    } else {
        I.grid_persist_ext = persistent_grid(k_extend<64, false>, cus);
        I.grid_persist_shadow = persistent_grid(k_shadow<64, false>, cus);
        I.grid_shade[kClsSubsurface] = persistent_grid(k_shade<kClsSubsurface, 64>, cus);
    }
    I.grid_shade[kClsTerminal] = persistent_grid(k_shade<kClsTerminal, 32>, cus);
    I.grid_shade[kClsDiffuse] = persistent_grid(k_shade<kClsDiffuse, 32>, cus);
    I.grid_shade[kClsPbr] = persistent_grid(k_shade<kClsPbr, 32>, cus);
    I.grid_shade[kClsDielectric] = persistent_grid(k_shade<kClsDielectric, 32>, cus);
    I.grid_linear = cus * 8;
    HIP_OK(hipDeviceSynchronize());
}

Renderer::~Renderer() {
    if (impl_) {
        for (auto& e : impl_->ev) {
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
        if (impl_->stream) (void)hipStreamDestroy(impl_->stream);
        delete impl_;
    }
}

int Renderer::width() const { return impl_->width; }
int Renderer::height() const { return impl_->height; }
int Renderer::batch() const { return impl_->batch; }
int Renderer::spp_done() const { return impl_->spp_done; }

// Stage timing with HIP events on the renderer's stream (profiled renderers only).
static hipEvent_t stage_begin(Renderer::Impl& I) {
    if (!I.profile) return nullptr;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, I.stream));
    return e;
}
static void stage_end(Renderer::Impl& I, int stage, hipEvent_t b) {
    if (!I.profile) return;
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    HIP_OK(hipEventRecord(e, I.stream));
    I.ev.push_back({stage, {b, e}});
}

template <int CLS, int STACK>
static void launch_shade(Renderer::Impl& I, const DevScene& sc, const FrameParams& fp, int q) {
    int* ctr = I.counters.as<int>();
    k_shade<CLS, STACK><<<I.grid_shade[CLS], kBlock, 0, I.stream>>>(
        sc, fp, I.qo[q].as<float4>(), I.qd[q].as<float4>(), I.qw[q].as<float4>(), I.hits.as<float4>(),
        I.lists.as<int>() + (size_t)CLS * I.cap, I.qo[q ^ 1].as<float4>(), I.qd[q ^ 1].as<float4>(),
        I.qw[q ^ 1].as<float4>(), I.shadow.as<float4>(), I.slot_d.as<float4>(), I.slot_i.as<float4>(), I.region, ctr,
        q);
    HIP_OK(hipGetLastError());
}

template <int STACK>
static void run_pass(Renderer::Impl& I, FrameParams fp, bool count) {
    const DevScene& sc = I.scene->dev();
    auto tick = [&]() { return stage_begin(I); };
    auto tock = [&](int stage, hipEvent_t e) { stage_end(I, stage, e); };
    int* ctr = I.counters.as<int>();
    unsigned long long* st = I.stats.as<unsigned long long>();
    const int cap = (int)I.cap;
    int q = 0;
    hipEvent_t b = tick();
    int grid_rg = (int)std::min<size_t>((fp.n_paths + kBlock - 1) / kBlock, (size_t)I.grid_linear * 4);
    k_raygen<<<grid_rg, kBlock, 0, I.stream>>>(fp, I.qo[q].as<float4>(), I.qd[q].as<float4>(), I.qw[q].as<float4>(),
                                               I.slot_d.as<float4>(), I.slot_i.as<float4>(), ctr, q, I.region);
    HIP_OK(hipGetLastError());
    tock(kStageRaygen, b);
    const GpuScene& gs = *I.scene;
    for (int bounce = 0; bounce < fp.max_bounces; ++bounce) {
        fp.bounce = bounce;
        b = tick();
        if (count)
            k_extend<STACK, true><<<I.grid_persist_ext, kBlock, 0, I.stream>>>(
                sc, I.qo[q].as<float4>(), I.qd[q].as<float4>(), I.hits.as<float4>(), I.lists.as<int>(), cap, I.region, ctr, q, st);
        else
            k_extend<STACK, false><<<I.grid_persist_ext, kBlock, 0, I.stream>>>(
                sc, I.qo[q].as<float4>(), I.qd[q].as<float4>(), I.hits.as<float4>(), I.lists.as<int>(), cap, I.region, ctr, q, st);
        HIP_OK(hipGetLastError());
        tock(kStageExtend, b);
        b = tick();
        // Classes absent from the scene get no launch; their lists stay empty.
        launch_shade<kClsTerminal, STACK>(I, sc, fp, q);
        if (gs.has_class(kClsDiffuse)) launch_shade<kClsDiffuse, STACK>(I, sc, fp, q);
        if (gs.has_class(kClsPbr)) launch_shade<kClsPbr, STACK>(I, sc, fp, q);
        if (gs.has_class(kClsDielectric)) launch_shade<kClsDielectric, STACK>(I, sc, fp, q);
        if (gs.has_class(kClsSubsurface)) launch_shade<kClsSubsurface, STACK>(I, sc, fp, q);
        tock(kStageShade, b);
        b = tick();
        if (count)
            k_shadow<STACK, true><<<I.grid_persist_shadow, kBlock, 0, I.stream>>>(
                sc, I.shadow.as<float4>(), I.slot_d.as<float4>(), I.slot_i.as<float4>(), I.region, ctr, st);
        else
            k_shadow<STACK, false><<<I.grid_persist_shadow, kBlock, 0, I.stream>>>(
                sc, I.shadow.as<float4>(), I.slot_d.as<float4>(), I.slot_i.as<float4>(), I.region, ctr, st);
        HIP_OK(hipGetLastError());
        tock(kStageShadow, b);
        q ^= 1;
    }
    b = tick();
    if (!I.set.flush_survivors)
        k_lost<<<I.grid_linear, kBlock, 0, I.stream>>>(I.qw[q].as<float4>(), I.slot_d.as<float4>(), I.slot_i.as<float4>(), ctr,
                                                       q, I.region);
    int k = fp.n_paths / I.wh;
    int grid_res = (int)std::min<size_t>((I.wh + kBlock - 1) / kBlock, (size_t)I.grid_linear * 4);
    k_resolve<<<grid_res, kBlock, 0, I.stream>>>(I.slot_d.as<float4>(), I.slot_i.as<float4>(), I.wh, k,
                                                 I.acc_d.as<float4>(), I.acc_i.as<float4>());
    HIP_OK(hipGetLastError());
    tock(kStageResolve, b);
}

void Renderer::render(int spp, bool count) {
    Impl& I = *impl_;
    int remaining = spp;
    while (remaining > 0) {
        int k = std::min(remaining, I.batch);
        FrameParams fp;
        fp.cam_pos = I.cam.position;
        fp.left_down = I.cam.left_down;
        fp.du = I.cam.du;
        fp.dv = I.cam.dv;
        fp.width = I.width;
        fp.height = I.height;
        fp.wh = I.wh;
        fp.seed = I.set.seed;
        fp.rr = I.set.rr;
        fp.rr_start = I.set.rr_start_depth;
        fp.max_bounces = I.set.max_bounces;
        fp.sample0 = I.sample_base + (uint32_t)I.spp_done;
        fp.n_paths = I.wh * k;
        fp.bounce = 0;
        if (I.scene->stack() == 32) run_pass<32>(I, fp, count);
        else run_pass<64>(I, fp, count);
        I.spp_done += k;
        remaining -= k;
    }
}

void Renderer::synchronize() { HIP_OK(hipStreamSynchronize(impl_->stream)); }

// ------------------------------------------------------------------ frames
void Renderer::set_camera(const CameraRec& c) {
    Impl& I = *impl_;
    if (c.width != I.width || c.height != I.height) throw std::invalid_argument("camera resolution differs from the renderer's");
    I.cam = c;
    reset();
}

const CameraRec& Renderer::camera() const { return impl_->cam; }
int Renderer::frames() const { return impl_->frames; }

void Renderer::reset_history() {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.stream));
    I.frames = 0;
}

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

static void ensure_frame_buffers(Renderer::Impl& I) {
    if (I.frame_alloc) return;
    size_t wh = (size_t)I.wh;
    for (int k = 0; k < 2; ++k) {
        I.g_pos[k].alloc(wh * 16);
        I.g_nrm[k].alloc(wh * 8);
        I.g_bary[k].alloc(wh * 8);
        I.g_md[k].alloc(wh * 16);
        I.h_d[k].alloc(wh * 16);
        I.h_i[k].alloc(wh * 16);
        I.m_d[k].alloc(wh * 8);
        I.m_i[k].alloc(wh * 8);
        I.h_len[k].alloc(wh * 4);
        I.f_d[k].alloc(wh * 16);
        I.f_i[k].alloc(wh * 16);
    }
    I.v_d.alloc(wh * 16);
    I.v_i.alloc(wh * 16);
    I.denoised.alloc(wh * 16);
    I.frame_alloc = true;
}

static GBufferView gview(Renderer::Impl& I, int k) {
    GBufferView v;
    v.pos_tri = I.g_pos[k].as<float4>();
    v.nrm_mat = I.g_nrm[k].as<ushort4>();
    v.bary_mesh = I.g_bary[k].as<ushort4>();
    v.motion_depth = I.g_md[k].as<float4>();
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

void Renderer::render_frame(int spp, const SvgfParams& p) {
    Impl& I = *impl_;
    if (spp < 1) throw std::invalid_argument("spp must be positive");
    ensure_frame_buffers(I);
    const int c = I.fcur, pv = c ^ 1;
    const bool first = I.frames == 0;
    const int w = I.width, h = I.height;
    const int tiles = ((w + 15) / 16) * ((h + 15) / 16);
    const DevScene& sc = I.scene->dev();

    hipEvent_t e = stage_begin(I);
    GBufferParams g = gbuffer_params(*I.host, I.cam, first ? I.cam : I.last_frame_cam);
    if (I.scene->stack() == 32) k_gbuffer<32><<<tiles, kBlock, 0, I.stream>>>(sc, g, gview(I, c));
    else k_gbuffer<64><<<tiles, kBlock, 0, I.stream>>>(sc, g, gview(I, c));
    HIP_OK(hipGetLastError());
    stage_end(I, kStageGBuffer, e);

    // Fresh samples of this frame only (the reference resets its frame buffer
    // every denoised frame: renderer.cuh:1297-1302, 1121-1124).
    k_zero4<<<I.grid_linear, kBlock, 0, I.stream>>>(I.acc_d.as<float4>(), (size_t)I.wh);
    k_zero4<<<I.grid_linear, kBlock, 0, I.stream>>>(I.acc_i.as<float4>(), (size_t)I.wh);
    HIP_OK(hipGetLastError());
    I.spp_done = 0;
    I.sample_base = (uint32_t)I.frames * (uint32_t)spp;
    render(spp);

    e = stage_begin(I);
    TemporalIO tio;
    tio.acc_d = I.acc_d.as<float4>();
    tio.acc_i = I.acc_i.as<float4>();
    tio.inv_spp = 1.f / (float)spp;
    tio.cur_d = I.h_d[c].as<float4>();
    tio.cur_i = I.h_i[c].as<float4>();
    tio.hist_d = I.h_d[pv].as<float4>();
    tio.hist_i = I.h_i[pv].as<float4>();
    tio.mom_prev_d = I.m_d[pv].as<float2>();
    tio.mom_prev_i = I.m_i[pv].as<float2>();
    tio.len_prev = I.h_len[pv].as<int>();
    tio.mom_d = I.m_d[c].as<float2>();
    tio.mom_i = I.m_i[c].as<float2>();
    tio.len = I.h_len[c].as<int>();
    k_temporal<<<tiles, kBlock, 0, I.stream>>>(w, h, first ? 1 : 0, p, tio, gview(I, c), gview(I, pv));
    HIP_OK(hipGetLastError());
    stage_end(I, kStageTemporal, e);

    e = stage_begin(I);
    FilterIO fio;
    fio.in_d = I.h_d[c].as<float4>();
    fio.in_i = I.h_i[c].as<float4>();
    fio.out_d = I.v_d.as<float4>();
    fio.out_i = I.v_i.as<float4>();
    fio.mom_d = I.m_d[c].as<float2>();
    fio.mom_i = I.m_i[c].as<float2>();
    fio.len = I.h_len[c].as<int>();
    k_variance<<<tiles, kBlock, 0, I.stream>>>(w, h, p, fio, gview(I, c));
    HIP_OK(hipGetLastError());
    stage_end(I, kStageVariance, e);

    I.out_sel = -1;
    if (!first) {
        // a-trous passes with step 1, 2, 4, ... (reference: renderer.cuh:1236-1250).
        int passes = std::max(1, std::min(8, p.atrous_iters));
        e = stage_begin(I);
        for (int k = 0; k < passes; ++k) {
            int o = k & 1;
            fio.in_d = k == 0 ? I.v_d.as<float4>() : I.f_d[o ^ 1].as<float4>();
            fio.in_i = k == 0 ? I.v_i.as<float4>() : I.f_i[o ^ 1].as<float4>();
            fio.out_d = I.f_d[o].as<float4>();
            fio.out_i = I.f_i[o].as<float4>();
            k_atrous<<<tiles, kBlock, 0, I.stream>>>(w, h, 1 << k, p, fio, gview(I, c));
            HIP_OK(hipGetLastError());
            I.out_sel = o;
        }
        stage_end(I, kStageAtrous, e);
    }
    const float4* od = I.out_sel < 0 ? I.v_d.as<float4>() : I.f_d[I.out_sel].as<float4>();
    const float4* oi = I.out_sel < 0 ? I.v_i.as<float4>() : I.f_i[I.out_sel].as<float4>();
    int grid_c = (int)std::min<size_t>((I.wh + kBlock - 1) / kBlock, (size_t)I.grid_linear * 4);
    k_combine<<<grid_c, kBlock, 0, I.stream>>>(od, oi, I.wh, I.denoised.as<float4>());
    HIP_OK(hipGetLastError());

    I.last_frame_cam = I.cam;
    I.flast = c;
    I.fcur = pv;
    I.frames++;
}

void Renderer::read_frame(int which, void* out) {
    Impl& I = *impl_;
    if (which < 0 || which >= kNumFrameBuffers) throw std::invalid_argument("unknown frame buffer");
    if (I.flast < 0) throw std::invalid_argument("no frame has been rendered");
    HIP_OK(hipStreamSynchronize(I.stream));
    const int c = I.flast;
    const DevBuf* src = nullptr;
    switch (which) {
        case kFrameDenoised: src = &I.denoised; break;
        case kFrameDirect: src = I.out_sel < 0 ? &I.v_d : &I.f_d[I.out_sel]; break;
        case kFrameIndirect: src = I.out_sel < 0 ? &I.v_i : &I.f_i[I.out_sel]; break;
        case kFrameHistoryDirect: src = &I.h_d[c]; break;
        case kFrameHistoryIndirect: src = &I.h_i[c]; break;
        case kFrameMomentsDirect: src = &I.m_d[c]; break;
        case kFrameMomentsIndirect: src = &I.m_i[c]; break;
        case kFrameHistoryLength: src = &I.h_len[c]; break;
        case kFrameVarianceDirect: src = &I.v_d; break;
        case kFrameVarianceIndirect: src = &I.v_i; break;
        case kFrameGPosTri: src = &I.g_pos[c]; break;
        case kFrameGNrmMat: src = &I.g_nrm[c]; break;
        case kFrameGBaryMesh: src = &I.g_bary[c]; break;
        case kFrameGMotionDepth: src = &I.g_md[c]; break;
    }
    HIP_OK(hipMemcpy(out, src->p, frame_buffer_bytes(which, I.width, I.height), hipMemcpyDeviceToHost));
}

void Renderer::reset() {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.stream));
    k_zero4<<<I.grid_linear, kBlock, 0, I.stream>>>(I.acc_d.as<float4>(), (size_t)I.wh);
    k_zero4<<<I.grid_linear, kBlock, 0, I.stream>>>(I.acc_i.as<float4>(), (size_t)I.wh);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(I.stream));
    I.spp_done = 0;
    I.sample_base = 0;
}

void Renderer::reset_stats() {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.stream));
    k_zero_u64<<<1, 64, 0, I.stream>>>(I.stats.as<unsigned long long>(), kNumStats);
    HIP_OK(hipStreamSynchronize(I.stream));
    for (auto& e : I.ev) {
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    I.ev.clear();
    for (int s = 0; s < kNumStages; ++s) {
        I.stage_ms[s] = 0;
        I.stage_launches[s] = 0;
    }
}

void Renderer::stats(RenderStats& out) {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.stream));
    unsigned long long s[kNumStats];
    HIP_OK(hipMemcpy(s, I.stats.p, sizeof(s), hipMemcpyDeviceToHost));
    for (auto& e : I.ev) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
        I.stage_ms[e.first] += ms;
        I.stage_launches[e.first] += 1;
        (void)hipEventDestroy(e.second.first);
        (void)hipEventDestroy(e.second.second);
    }
    I.ev.clear();
    out.ext_rays = s[kStatExtRays];
    out.shadow_rays = s[kStatShadowRays];
    out.ext_nodes = s[kStatExtNodes];
    out.ext_tris = s[kStatExtTris];
    out.shadow_nodes = s[kStatShadowNodes];
    out.shadow_tris = s[kStatShadowTris];
    out.probe_rays = s[kStatProbeRays];
    for (int k = 0; k < kNumStages; ++k) {
        out.stage_ms[k] = I.stage_ms[k];
        out.stage_launches[k] = I.stage_launches[k];
    }
    out.spp_done = I.spp_done;
    out.batch_spp = I.batch;
}

void Renderer::read_accum(int which, float* out) {
    Impl& I = *impl_;
    HIP_OK(hipStreamSynchronize(I.stream));
    std::vector<float> d((size_t)I.wh * 4), ind((size_t)I.wh * 4);
    HIP_OK(hipMemcpy(d.data(), I.acc_d.p, d.size() * 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ind.data(), I.acc_i.p, ind.size() * 4, hipMemcpyDeviceToHost));
    size_t n = (size_t)I.wh * 4;
    if (which == 0) std::memcpy(out, d.data(), n * 4);
    else if (which == 1) std::memcpy(out, ind.data(), n * 4);
    else {
        int spp = I.spp_done > 0 ? I.spp_done : 1;
        for (size_t k = 0; k < (size_t)I.wh; ++k) {
            for (int c = 0; c < 3; ++c) out[4 * k + c] = (d[4 * k + c] + ind[4 * k + c]) / (float)spp;
            out[4 * k + 3] = 1.f;
        }
    }
}

// ------------------------------------------------------------------ tools
void trace_batch(const GpuScene& scene, const float* rays, int n, float* out_hits, uint32_t* out_counts) {
    DevBuf r, o, c;
    r.alloc((size_t)n * 32);
    o.alloc((size_t)n * 16);
    c.alloc((size_t)n * 8);
    HIP_OK(hipMemcpy(r.p, rays, (size_t)n * 32, hipMemcpyHostToDevice));
    int grid = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
    if (scene.stack() == 32)
        k_trace_batch<32><<<grid, kBlock>>>(scene.dev(), r.as<float4>(), n, o.as<float4>(), c.as<uint2>());
    else
        k_trace_batch<64><<<grid, kBlock>>>(scene.dev(), r.as<float4>(), n, o.as<float4>(), c.as<uint2>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(out_hits, o.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (out_counts) HIP_OK(hipMemcpy(out_counts, c.p, (size_t)n * 8, hipMemcpyDeviceToHost));
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
