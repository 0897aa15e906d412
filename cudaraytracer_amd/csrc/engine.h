// Wavefront engine internals shared by the single-scene renderer and the
// multi-job pool: device buffers, stage timing, render passes over one or
// many jobs, and the per-view G-buffer / SVGF state.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "render.h"

namespace rt {

#define HIP_OK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                     #x);                                                           \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
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

// Suspended extension searches: room for one in kContPathsPer paths, each a
// float4 hit, an int4 state and kContStack stack words.
constexpr size_t kContPathsPer = 4;
constexpr size_t kContStack = 16;
constexpr size_t kContBytes = 16 + 16 + 4 * kContStack;
// Device bytes per path of an engine's queues: two ping-pong SoA queues
// (3 x float4 each), the hit record, one class-list entry per class, the
// shadow record, the per-sample direct/indirect sums and the share of the
// suspended-search records.
constexpr size_t kQueueBytesPerPath = 2 * 3 * 16 + 16 + 4 * kNumClasses + 64 + 32 + kContBytes / kContPathsPer;
// With the ray sort on: a 2-byte key, a 4-byte permutation entry and a
// 32-byte copy of the ray at its sorted position per path.
constexpr size_t kSortBytesPerPath = 38;
// Device bytes per pixel of a FrameState: two G-buffers (48 B each), two
// histories (2 x float4), moments (2 x float2) and lengths (int) per frame
// parity, the variance output, the a-trous ping-pong pair, the denoised colour,
// the filter guide (decoded normal + depth, key + history length) and the
// luminance guide of the variance pass's input (float4).  (The variance
// pass's tile list adds under 1/16 B per pixel plus a row of ints.)
constexpr size_t kFrameStateBytesPerPixel = 2 * 48 + 2 * 32 + 2 * 16 + 2 * 4 + 32 + 2 * 32 + 16 + 24 + 16;

// Per-stage GPU time from HIP events recorded on the work's own stream.
struct Profiler {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    double ms[kNumStages] = {0};
    int launches[kNumStages] = {0};
    hipEvent_t begin(hipStream_t s);
    void end(int stage, hipEvent_t b, hipStream_t s);
    void collect();  // after the stream has been synchronised
    void reset();
    ~Profiler();
};

// One job's share of a render pass: `spp` samples from `fp.sample0` of every
// pixel, summed in sample order into the job's accumulators.
struct PassItem {
    const GpuScene* scene = nullptr;
    FrameParams fp{};
    int spp = 0;
    int flush_survivors = 0;
    int env_alias = 0;
    float4* acc_d = nullptr;
    float4* acc_i = nullptr;
};

// Path queues, counters and the pass driver.  A pass holds at most
// capacity() paths and kMaxPassJobs jobs.
class Engine {
public:
    Engine(size_t capacity_paths, bool profile);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    size_t capacity() const { return cap_; }
    hipStream_t stream() const { return stream_; }
    // Enqueue one pass on the engine's stream (asynchronous).
    void run(const std::vector<PassItem>& items, bool count);
    void read_stats(unsigned long long out[kNumStats]);
    void reset_stats();
    Profiler prof;
    // Replay passes as captured HIP graphs (off under profiling or counting).
    void set_use_graphs(bool on) { use_graphs_ = on; }
    // Persistent stage grids = occupancy x CUs x scale (a tuning knob).
    void set_grid_scale(float scale);
    // Reorder each bounce's extension rays by direction octant and origin
    // cell before extend (kSortBytesPerPath more device memory per path).
    void set_ray_sort(bool on);
    // Shadow rays stop at the first hit clearly in front of the shading point
    // (the any-hit fast mode) instead of finding the closest hit.
    // Shadow test: 0 the reference's, 1 its early-exit form, 2 the segment test (k_shadow).
    void set_shadow_mode(int mode) { shadow_mode_ = mode; }
    float grid_scale() const { return grid_scale_; }

private:
    template <bool POOLED>
    void run_t(int max_bounces, const bool* classes, bool count, int stack_entries, bool suspend,
               bool suspend_shadow);
    struct Grids {
        int ext = 0, ext_susp = 0, ext_cont = 0, shadow = 0, shadow_susp = 0, shadow_cont = 0, shadow_early = 0, shadow_segment = 0, shade = 0, shade_stack = 0;
    };
    template <bool POOLED>
    Grids grids_for(int stack_entries);
    std::map<uint64_t, Grids> grids_;
    int cus_ = 0;
    bool use_graphs_ = true;
    float grid_scale_ = 1.f;
    std::map<uint64_t, hipGraphExec_t> graphs_;
    size_t cap_ = 0;
    hipStream_t stream_ = nullptr;
    DevBuf qo_[2], qd_[2], qw_[2], hits_, lists_, shadow_, slot_d_, slot_i_, counters_, stats_;
    DevBuf cont_hit_, cont_state_, cont_stack_;  // suspended extension searches (k_extend_cont)
    int cont_room_ = 0;                          // records per slice
    bool ray_sort_ = false;
    int shadow_mode_ = 0;
    DevBuf keys_, perm_, sort_hist_, sort_cur_;  // ray sort: key and permutation per path, per-slice bins
    DevBuf sorted_o_, sorted_d_;                 // ray sort: extend's copy of each ray at its sorted position
    // Pass tables: one device table, filled by stream-ordered copies from a
    // ring of pinned host tables.
    static constexpr int kRing = 8;
    DevBuf tables_;
    void* host_tables_ = nullptr;
    hipEvent_t ring_ev_[kRing] = {};
    int ring_next_ = 0;
    int grid_linear_ = 0;
};

// G-buffers and SVGF state of one width x height view (reference: the
// G-buffer pair renderer.cuh:1356-1503 and the denoise buffers 1309-1355).
class FrameState {
public:
    FrameState(int width, int height);
    // Start a frame seen from `cam`: G-buffer against the previous frame's
    // camera, and cleared accumulators for the frame's samples.  With a
    // G-buffer stream the G-buffer runs there, beside the frame's path tracing
    // on `s` (it needs only the camera); `s` waits for it in end().
    void begin(const GpuScene& gs, const HostScene& host, const CameraRec& cam, float4* acc_d, float4* acc_i,
               Profiler& prof, hipStream_t s, hipStream_t gbuf_stream = nullptr);
    ~FrameState();
    FrameState(const FrameState&) = delete;
    FrameState& operator=(const FrameState&) = delete;
    // Finish it: temporal blend of the frame's `spp` samples, variance, a-trous, combine.
    void end(int spp, const SvgfParams& p, const float4* acc_d, const float4* acc_i, Profiler& prof,
             hipStream_t s);
    void reset_history() { frames_ = 0; }
    // Make the state reusable for a new job: no history, nothing readable.
    void recycle() {
        frames_ = 0;
        fcur_ = 0;
        flast_ = -1;
        out_sel_ = -1;
        started_ = false;
    }
    int width() const { return w_; }
    int height() const { return h_; }
    // Tile shape of the per-pixel kernels: 2^shift x (256 >> shift), shift in 3..7.
    void set_tile_shift(int shift) {
        if (shift < 3 || shift > 7) throw std::invalid_argument("tile shift must be in 3..7");
        tile_shift_ = shift;
    }
    int frames() const { return frames_; }
    // Sample index of the next frame's first sample.
    uint32_t next_sample0(int spp) const { return (uint32_t)frames_ * (uint32_t)spp; }
    void read(int which, void* out, hipStream_t s);
    size_t device_bytes() const;

private:
    int w_, h_, wh_;
    DevBuf g_pos_[2], g_nrm_[2], g_bary_[2], g_md_[2];
    DevBuf h_d_[2], h_i_[2], m_d_[2], m_i_[2], h_len_[2];
    DevBuf v_d_, v_i_, f_d_[2], f_i_[2], denoised_;
    DevBuf guide_nd_, guide_kl_;  // this frame's filter guide
    DevBuf lum_;                  // luminance guide of the variance pass's input
    DevBuf short_tiles_, short_n_;  // tiles holding a short history; two sets of per-band counts
    int short_sel_ = 0;             // which set of counts the next temporal pass fills
    CameraRec last_cam_;
    int fcur_ = 0, flast_ = -1, frames_ = 0, out_sel_ = -1;
    int tile_shift_ = 4;
    bool started_ = false;
    hipEvent_t ev_order_ = nullptr, ev_gbuf_ = nullptr;  // G-buffer stream ordering
    bool gbuf_pending_ = false;                          // end() must wait for ev_gbuf_
};

// Read accumulators (which: 0 direct sums, 1 indirect sums, 2 (direct + indirect) / spp).
void read_accum_buffers(const float4* acc_d, const float4* acc_i, int wh, int spp_done, int which, float* out,
                        hipStream_t s);

// FrameParams of a camera for a job's pass.
FrameParams frame_params(const CameraRec& cam, const RenderSettings& set, uint32_t sample0, int spp);

}  // namespace rt
