// Multi-job scheduling over one wavefront engine, and the arrival-driven
// serving loop (continuous batching vs. the plain one-job-at-a-time loop).
//
// The reference renders one scene per process in an interactive loop
// (Src/main.cu:26-33, Src/Core/renderer.cuh:1106-1307); a pool keeps many
// scenes resident and carries the next work of many jobs in every pass.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <thread>

#include "engine.h"

namespace rt {

namespace {
using Clock = std::chrono::steady_clock;
}

struct Pool::Impl {
    PoolSettings set;
    std::unique_ptr<Engine> eng;
    struct SceneSlot {
        const HostScene* host;
        const GpuScene* gpu;
    };
    std::vector<SceneSlot> scenes;
    struct Job {
        int id = 0;
        JobDesc d;
        JobStatus st;
        int wh = 0;
        std::unique_ptr<DevBuf> acc_d, acc_i;
        std::unique_ptr<FrameState> fs;
        size_t bytes = 0;
        // Pipelined real-time jobs: cohort (which ticks trace it) and phase
        // (0 = frame not begun, 1 = G-buffer done, ready to trace, 2 = traced,
        // SVGF pending).
        int cohort = 0, phase = 0;
    };
    // Device state of released jobs, reused by later jobs of the same size
    // (admission then needs no allocation).
    std::multimap<int, std::unique_ptr<DevBuf>> spare_acc;                       // by W*H
    std::multimap<std::pair<int, int>, std::unique_ptr<FrameState>> spare_fs;    // by (W, H)
    size_t spare_bytes = 0;
    void free_spares() {
        spare_acc.clear();
        spare_fs.clear();
        spare_bytes = 0;
    }
    std::map<int, std::unique_ptr<Job>> jobs;
    std::deque<int> active;  // submission order
    hipStream_t side = nullptr;  // frame work (G-buffer, SVGF) beside the engine's pass
    // A third stream for the serial tick's filters: the frames of a tick are
    // filtered on the engine stream, the side stream and this one in turn, so
    // one frame's kernels fill the tails of another's.
    hipStream_t filt = nullptr;
    hipEvent_t ev_pass = nullptr, ev_side = nullptr, ev_filt = nullptr;
    Pool::Tuning tune;
    long tick = 0;
    int cohort_size[2] = {0, 0};
    int next_id = 0;
    size_t scene_bytes = 0, job_bytes = 0, hbm_budget = 0;
    Clock::time_point t0;

    Job& job(int id) const {
        auto it = jobs.find(id);
        if (it == jobs.end()) throw std::invalid_argument("unknown job id");
        return *it->second;
    }
};

Pool::Pool(const PoolSettings& s) : impl_(new Impl) {
    Impl& I = *impl_;
    I.set = s;
    size_t qb = s.queue_budget_bytes ? s.queue_budget_bytes : ((size_t)8 << 30);
    I.eng.reset(new Engine(std::max<size_t>(1, qb / kQueueBytesPerPath), s.profile != 0));
    I.hbm_budget = s.hbm_budget_bytes ? s.hbm_budget_bytes : ((size_t)64 << 30);
    HIP_OK(hipStreamCreateWithFlags(&I.side, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&I.filt, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&I.ev_pass, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&I.ev_side, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&I.ev_filt, hipEventDisableTiming));
    I.t0 = Clock::now();
}

Pool::~Pool() {
    if (impl_) {
        if (impl_->eng) (void)hipStreamSynchronize(impl_->eng->stream());
        if (impl_->side) (void)hipStreamSynchronize(impl_->side);
        if (impl_->filt) (void)hipStreamSynchronize(impl_->filt);
        hipStream_t side = impl_->side, filt = impl_->filt;
        hipEvent_t evs[3] = {impl_->ev_pass, impl_->ev_side, impl_->ev_filt};
        delete impl_;
        if (side) (void)hipStreamDestroy(side);
        if (filt) (void)hipStreamDestroy(filt);
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
    }
}

double Pool::now() const { return std::chrono::duration<double>(Clock::now() - impl_->t0).count(); }
size_t Pool::capacity() const { return impl_->eng->capacity(); }
size_t Pool::bytes_in_use() const { return impl_->scene_bytes + impl_->job_bytes; }
bool Pool::idle() const { return impl_->active.empty(); }

int Pool::add_scene(const HostScene& host, const GpuScene& gpu) {
    Impl& I = *impl_;
    I.scenes.push_back({&host, &gpu});
    I.scene_bytes += gpu.device_bytes();
    return (int)I.scenes.size() - 1;
}

const HostScene& Pool::scene_host(int slot) const {
    if (slot < 0 || slot >= (int)impl_->scenes.size()) throw std::invalid_argument("unknown scene slot");
    return *impl_->scenes[slot].host;
}

const GpuScene& Pool::scene_gpu(int slot) const {
    if (slot < 0 || slot >= (int)impl_->scenes.size()) throw std::invalid_argument("unknown scene slot");
    return *impl_->scenes[slot].gpu;
}

size_t Pool::check(const JobDesc& d) const {
    const Impl& I = *impl_;
    if (d.scene < 0 || d.scene >= (int)I.scenes.size()) throw std::invalid_argument("unknown scene slot");
    if (d.spp < 1) throw std::invalid_argument("spp must be positive");
    if (d.frames < 0) throw std::invalid_argument("frames must be >= 0");
    if (d.frames > 0 && (int)d.cams.size() != d.frames) throw std::invalid_argument("one camera per frame is required");
    if (d.frames == 0 && d.cams.size() > 1) throw std::invalid_argument("an accumulate job takes at most one camera");
    if (d.set.max_bounces < 0 || d.set.max_bounces > 255) throw std::invalid_argument("max_bounces out of range");
    const HostScene& h = *I.scenes[d.scene].host;
    const CameraRec& c0 = d.cams.empty() ? h.camera : d.cams[0];
    for (const auto& c : d.cams)
        if (c.width != c0.width || c.height != c0.height) throw std::invalid_argument("cameras of a job differ in size");
    const size_t wh = (size_t)c0.width * c0.height;
    if (wh * (size_t)(d.frames > 0 ? d.spp : 1) > I.eng->capacity())
        throw std::invalid_argument("one frame of the job exceeds the pool's path capacity");
    size_t bytes = wh * 32 + (d.frames > 0 ? wh * kFrameStateBytesPerPixel : 0);
    if (I.scene_bytes + bytes > I.hbm_budget) throw std::invalid_argument("the job alone exceeds the HBM budget");
    return bytes;
}

int Pool::submit(const JobDesc& d) {
    Impl& I = *impl_;
    const size_t bytes = check(d);
    const HostScene& h = *I.scenes[d.scene].host;
    const CameraRec& c0 = d.cams.empty() ? h.camera : d.cams[0];
    const size_t wh = (size_t)c0.width * c0.height;
    // Bytes this job takes from spare state of the same size, and bytes it allocates.
    size_t from_spare = 0;
    if (I.spare_acc.count((int)wh) >= 2) from_spare += wh * 32;
    if (d.frames > 0 && I.spare_fs.count({c0.width, c0.height}) > 0) from_spare += wh * kFrameStateBytesPerPixel;
    if (I.scene_bytes + I.job_bytes + I.spare_bytes + (bytes - from_spare) > I.hbm_budget) {
        I.free_spares();  // spare state of other sizes gives way first
        from_spare = 0;
    }
    if (I.scene_bytes + I.job_bytes + I.spare_bytes + (bytes - from_spare) > I.hbm_budget) return -1;
    std::unique_ptr<Impl::Job> j(new Impl::Job);
    j->id = I.next_id++;
    j->d = d;
    if (j->d.cams.empty()) j->d.cams.push_back(h.camera);
    j->wh = (int)wh;
    j->bytes = bytes;
    auto take_acc = [&]() {
        auto it = I.spare_acc.find((int)wh);
        if (it != I.spare_acc.end()) {
            std::unique_ptr<DevBuf> b = std::move(it->second);
            I.spare_acc.erase(it);
            I.spare_bytes -= wh * 16;
            return b;
        }
        std::unique_ptr<DevBuf> b(new DevBuf);
        b->alloc(wh * 16);
        return b;
    };
    j->acc_d = take_acc();
    j->acc_i = take_acc();
    HIP_OK(hipMemsetAsync(j->acc_d->p, 0, wh * 16, I.eng->stream()));
    HIP_OK(hipMemsetAsync(j->acc_i->p, 0, wh * 16, I.eng->stream()));
    if (d.frames > 0) {
        auto it = I.spare_fs.find({c0.width, c0.height});
        if (it != I.spare_fs.end()) {
            j->fs = std::move(it->second);
            I.spare_fs.erase(it);
            I.spare_bytes -= wh * kFrameStateBytesPerPixel;
            j->fs->recycle();
        } else {
            j->fs.reset(new FrameState(c0.width, c0.height));
        }
        j->fs->set_tile_shift(I.tune.tile_shift);
    }
    if (d.frames > 0) {
        // The cohort whose frame work runs in the next tick begins soonest.
        int first = (int)((I.tick & 1) ^ 1);
        j->cohort = I.cohort_size[first] <= I.cohort_size[first ^ 1] ? first : first ^ 1;
        I.cohort_size[j->cohort]++;
        j->phase = 0;
    }
    j->st.submit_s = now();
    I.job_bytes += bytes;
    int id = j->id;
    I.active.push_back(id);
    I.jobs[id] = std::move(j);
    return id;
}

namespace {
// Paths per pass: the tuned limit, but never below the largest single unit of
// work of an active job (one frame, or one sample of every pixel), so every
// job can be scheduled.
size_t pass_limit(Pool::Impl& I) {
    size_t cap = I.eng->capacity();
    if (!I.tune.pass_paths || I.tune.pass_paths >= cap) return cap;
    size_t unit = 0;
    for (int id : I.active) {
        const Pool::Impl::Job& j = I.job(id);
        unit = std::max(unit, (size_t)j.wh * (size_t)(j.d.frames > 0 ? j.d.spp : 1));
    }
    return std::min(cap, std::max(I.tune.pass_paths, unit));
}

PassItem pass_item(const GpuScene* gpu, const JobDesc& d, const CameraRec& cam, uint32_t sample0, int k,
                   float4* acc_d, float4* acc_i) {
    PassItem it;
    it.scene = gpu;
    it.fp = frame_params(cam, d.set, sample0, k);
    it.spp = k;
    it.flush_survivors = d.set.flush_survivors;
    it.env_alias = d.set.env_alias;
    it.acc_d = acc_d;
    it.acc_i = acc_i;
    return it;
}
}  // namespace

int Pool::max_realtime() const { return impl_->set.max_realtime; }

int Pool::step() {
    Impl& I = *impl_;
    if (I.active.empty()) return 0;
    if (I.set.pipeline) {
        for (int guard = 0; guard < 4; ++guard) {
            int n = step_pipelined();
            if (n > 0) return n;
        }
        throw std::runtime_error("no active job fits the pool's path capacity");
    }
    return step_serial();
}

// Two cohorts of real-time jobs alternate: in a tick the engine's stream
// traces one cohort's frames (and every accumulate job's next samples) while
// the side stream finishes the other cohort's previous frames (temporal,
// variance, a-trous) and starts their next ones (G-buffer, cleared
// accumulators).  Path tracing (latency-bound traversal) and the filters
// (ALU-bound) overlap on the device; each job still sees its frames in order.
int Pool::step_pipelined() {
    Impl& I = *impl_;
    hipStream_t s = I.eng->stream();
    const int c = (int)(I.tick & 1), o = c ^ 1;
    const size_t cap = pass_limit(I);
    std::vector<PassItem> items;
    std::vector<Impl::Job*> traced;
    std::vector<int> spp_of;
    std::vector<Impl::Job*> side_jobs;
    size_t used = 0;
    const double t_start = now();
    for (int id : I.active) {
        Impl::Job& j = I.job(id);
        if (j.d.frames > 0 && j.cohort == o) {
            if (j.phase == 2 || j.phase == 0) side_jobs.push_back(&j);
            continue;
        }
        if (j.d.frames > 0 && j.phase == 0) {
            side_jobs.push_back(&j);  // a cohort-c job not begun yet: begin now, trace on its next turn
            continue;
        }
        if ((int)items.size() == kMaxPassJobs) continue;
        if (j.d.frames > 0 && j.phase != 1) continue;
        const Impl::SceneSlot& sc = I.scenes[j.d.scene];
        int k;
        uint32_t sample0;
        const CameraRec* cam;
        if (j.d.frames > 0) {
            k = j.d.spp;
            if (used + (size_t)j.wh * k > cap) continue;
            sample0 = j.fs->next_sample0(k);
            cam = &j.d.cams[j.st.frames_done];
        } else {
            int remaining = j.d.spp - j.st.spp_done;
            size_t fit = (cap - used) / (size_t)j.wh;
            k = (int)std::min<size_t>((size_t)remaining, fit);
            if (k < 1) continue;
            sample0 = (uint32_t)j.st.spp_done;
            cam = &j.d.cams[0];
        }
        items.push_back(pass_item(sc.gpu, j.d, *cam, sample0, k, j.acc_d->as<float4>(), j.acc_i->as<float4>()));
        traced.push_back(&j);
        spp_of.push_back(k);
        used += (size_t)j.wh * k;
    }
    // Side streams: finish traced frames of cohort o, begin the next ones; a
    // job's frame work stays on one stream, the jobs alternate between two.
    hipStream_t sides[2] = {I.side, I.filt};
    int ns = 0;
    for (Impl::Job* j : side_jobs) {
        const Impl::SceneSlot& sc = I.scenes[j->d.scene];
        hipStream_t ss = sides[ns++ & 1];
        if (j->phase == 2) {
            j->fs->end(j->d.spp, j->d.svgf, j->acc_d->as<float4>(), j->acc_i->as<float4>(), I.eng->prof, ss);
            if (j->st.frames_done + 1 == j->d.frames) continue;
            j->fs->begin(*sc.gpu, *sc.host, j->d.cams[j->st.frames_done + 1], j->acc_d->as<float4>(),
                         j->acc_i->as<float4>(), I.eng->prof, ss);
        } else {
            if (j->st.start_s < 0) j->st.start_s = t_start;
            j->fs->begin(*sc.gpu, *sc.host, j->d.cams[j->st.frames_done], j->acc_d->as<float4>(),
                         j->acc_i->as<float4>(), I.eng->prof, ss);
        }
    }
    if (!items.empty()) I.eng->run(items, false);
    HIP_OK(hipStreamSynchronize(I.side));
    HIP_OK(hipStreamSynchronize(I.filt));
    HIP_OK(hipStreamSynchronize(s));
    const double t_done = now();
    I.tick++;
    int served = 0;
    for (Impl::Job* j : side_jobs) {
        if (j->phase == 2) {
            j->st.frames_done++;
            j->st.spp_done = j->d.spp;
            j->st.frame_done_s.push_back(t_done);
            if (j->st.frames_done == j->d.frames) {
                j->st.done = 1;
                j->st.done_s = t_done;
                I.cohort_size[j->cohort]--;
                I.active.erase(std::find(I.active.begin(), I.active.end(), j->id));
            }
        }
        j->phase = 1;
        ++served;
    }
    for (size_t k = 0; k < traced.size(); ++k) {
        Impl::Job* j = traced[k];
        if (j->st.start_s < 0) j->st.start_s = t_start;
        if (j->d.frames > 0) {
            j->phase = 2;
        } else {
            j->st.spp_done += spp_of[k];
            if (j->st.spp_done == j->d.spp) {
                j->st.done = 1;
                j->st.done_s = t_done;
                I.active.erase(std::find(I.active.begin(), I.active.end(), j->id));
            }
        }
        ++served;
    }
    return served;
}

int Pool::step_serial() {
    Impl& I = *impl_;
    if (I.active.empty()) return 0;
    hipStream_t s = I.eng->stream();
    const size_t cap = pass_limit(I);
    std::vector<PassItem> items;
    std::vector<Impl::Job*> served;
    std::vector<int> spp_of;
    size_t used = 0;
    for (int id : I.active) {
        if ((int)items.size() == kMaxPassJobs) break;
        Impl::Job& j = I.job(id);
        const Impl::SceneSlot& sc = I.scenes[j.d.scene];
        int k;
        if (j.d.frames > 0) {
            k = j.d.spp;
            if (used + (size_t)j.wh * k > cap) continue;
        } else {
            int remaining = j.d.spp - j.st.spp_done;
            size_t fit = (cap - used) / (size_t)j.wh;
            k = (int)std::min<size_t>((size_t)remaining, fit);
            if (k < 1) continue;
        }
        PassItem it;
        it.scene = sc.gpu;
        const CameraRec& cam = j.d.frames > 0 ? j.d.cams[j.st.frames_done] : j.d.cams[0];
        uint32_t sample0 = j.d.frames > 0 ? j.fs->next_sample0(k) : (uint32_t)j.st.spp_done;
        it.fp = frame_params(cam, j.d.set, sample0, k);
        it.spp = k;
        it.flush_survivors = j.d.set.flush_survivors;
        it.env_alias = j.d.set.env_alias;
        it.acc_d = j.acc_d->as<float4>();
        it.acc_i = j.acc_i->as<float4>();
        items.push_back(it);
        served.push_back(&j);
        spp_of.push_back(k);
        used += (size_t)j.wh * k;
    }
    if (items.empty()) throw std::runtime_error("no active job fits the pool's path capacity");
    const double t_start = now();
    for (Impl::Job* j : served) {
        if (j->st.start_s < 0) j->st.start_s = t_start;
        if (j->d.frames > 0)
            j->fs->begin(*I.scenes[j->d.scene].gpu, *I.scenes[j->d.scene].host, j->d.cams[j->st.frames_done],
                         j->acc_d->as<float4>(), j->acc_i->as<float4>(), I.eng->prof, s, I.side);
    }
    I.eng->run(items, false);
    // Each frame's filters (temporal, variance, a-trous) read only its own job's
    // buffers: the frames go to three streams in turn, after the pass, and the
    // engine stream waits for all of them (the next tick's begin clears the
    // accumulators the filters read).
    hipStream_t fstreams[3] = {s, I.side, I.filt};
    bool forked = false;
    int nf = 0;
    for (size_t k = 0; k < served.size(); ++k) {
        Impl::Job* j = served[k];
        if (j->d.frames <= 0) continue;
        hipStream_t fs = fstreams[nf++ % 3];
        if (fs != s && !forked) {
            HIP_OK(hipEventRecord(I.ev_pass, s));
            HIP_OK(hipStreamWaitEvent(I.side, I.ev_pass, 0));
            HIP_OK(hipStreamWaitEvent(I.filt, I.ev_pass, 0));
            forked = true;
        }
        j->fs->end(spp_of[k], j->d.svgf, j->acc_d->as<float4>(), j->acc_i->as<float4>(), I.eng->prof, fs);
    }
    if (forked) {
        HIP_OK(hipEventRecord(I.ev_side, I.side));
        HIP_OK(hipEventRecord(I.ev_filt, I.filt));
        HIP_OK(hipStreamWaitEvent(s, I.ev_side, 0));
        HIP_OK(hipStreamWaitEvent(s, I.ev_filt, 0));
    }
    HIP_OK(hipStreamSynchronize(s));
    const double t_done = now();
    for (size_t k = 0; k < served.size(); ++k) {
        Impl::Job* j = served[k];
        if (j->d.frames > 0) {
            j->st.frames_done++;
            j->st.spp_done = spp_of[k];
            j->st.frame_done_s.push_back(t_done);
            if (j->st.frames_done == j->d.frames) j->st.done = 1;
        } else {
            j->st.spp_done += spp_of[k];
            if (j->st.spp_done == j->d.spp) j->st.done = 1;
        }
        if (j->st.done) {
            j->st.done_s = t_done;
            I.active.erase(std::find(I.active.begin(), I.active.end(), j->id));
        }
    }
    return (int)served.size();
}

const JobStatus& Pool::status(int job) const { return impl_->job(job).st; }

void Pool::job_size(int job, int& width, int& height) const {
    const Impl::Job& j = impl_->job(job);
    width = j.d.cams[0].width;
    height = j.d.cams[0].height;
}

void Pool::read_accum(int job, int which, float* out) {
    Impl::Job& j = impl_->job(job);
    read_accum_buffers(j.acc_d->as<float4>(), j.acc_i->as<float4>(), j.wh, j.st.spp_done, which, out,
                       impl_->eng->stream());
}

void Pool::read_frame(int job, int which, void* out) {
    Impl::Job& j = impl_->job(job);
    if (!j.fs) throw std::invalid_argument("not a real-time job");
    j.fs->read(which, out, impl_->eng->stream());
}

void Pool::release(int job) {
    Impl& I = *impl_;
    Impl::Job& j = I.job(job);
    HIP_OK(hipStreamSynchronize(I.eng->stream()));
    HIP_OK(hipStreamSynchronize(I.side));
    auto it = std::find(I.active.begin(), I.active.end(), job);
    if (it != I.active.end()) I.active.erase(it);
    if (j.fs && !j.st.done) I.cohort_size[j.cohort]--;
    I.job_bytes -= j.bytes;
    I.spare_bytes += (size_t)j.wh * 32;
    I.spare_acc.emplace(j.wh, std::move(j.acc_d));
    I.spare_acc.emplace(j.wh, std::move(j.acc_i));
    if (j.fs) {
        I.spare_bytes += (size_t)j.wh * kFrameStateBytesPerPixel;
        std::pair<int, int> key(j.fs->width(), j.fs->height());
        I.spare_fs.emplace(key, std::move(j.fs));
    }
    I.jobs.erase(job);
}

void Pool::stats(RenderStats& out) {
    Impl& I = *impl_;
    unsigned long long s[kNumStats];
    I.eng->read_stats(s);
    std::memset(&out, 0, sizeof(out));
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
}

void Pool::reset_stats() { impl_->eng->reset_stats(); }

void Pool::set_tuning(const Tuning& t) {
    Impl& I = *impl_;
    if (t.tile_shift < 3 || t.tile_shift > 7) throw std::invalid_argument("tile shift must be in 3..7");
    if (t.ray_sort != 0 && t.ray_sort != 1) throw std::invalid_argument("ray_sort must be 0 or 1");
    I.eng->set_grid_scale(t.grid_scale);
    I.eng->set_ray_sort(t.ray_sort != 0);
    I.tune = t;
    for (auto& kv : I.jobs)
        if (kv.second->fs) kv.second->fs->set_tile_shift(t.tile_shift);
}

Pool::Tuning Pool::tuning() const { return impl_->tune; }

Pool::TuneReport Pool::autotune(const std::vector<JobDesc>& sample) {
    Impl& I = *impl_;
    if (!I.active.empty()) throw std::invalid_argument("autotune needs an idle pool");
    if (sample.empty()) throw std::invalid_argument("autotune needs sample jobs");
    size_t sample_paths = 0;
    bool frames = false;
    for (const auto& d : sample) {
        check(d);
        const CameraRec& c = d.cams.empty() ? I.scenes[d.scene].host->camera : d.cams[0];
        sample_paths += (size_t)c.width * c.height * d.spp * (size_t)std::max(1, d.frames);
        frames = frames || d.frames > 0;
    }
    TuneReport rep;
    auto run = [&](const Tuning& t) {
        set_tuning(t);
        double best = 1e30;
        for (int rep_i = 0; rep_i < 2; ++rep_i) {
            HIP_OK(hipDeviceSynchronize());
            const double t0 = now();
            std::vector<int> ids;
            for (const auto& d : sample) {
                int id = submit(d);
                if (id < 0) throw std::invalid_argument("the autotune sample does not fit the HBM budget");
                ids.push_back(id);
            }
            while (step() > 0) {
            }
            best = std::min(best, (now() - t0) * 1e3);
            for (int id : ids) release(id);
            rep.runs++;
        }
        return best;
    };
    Tuning cur = I.tune;
    run(cur);  // warm-up: kernel loading, graph capture, state allocation
    double best_ms = run(cur);
    rep.baseline_ms = best_ms;
    auto try_values = [&](auto setter, const auto& values) {
        Tuning keep = cur;
        for (const auto& v : values) {
            Tuning t = cur;
            setter(t, v);
            double ms = run(t);
            if (ms < best_ms * 0.98) {  // a change has to win by 2 %
                best_ms = ms;
                keep = t;
            }
        }
        cur = keep;
    };
    const float scales[] = {0.5f, 0.75f, 1.25f};
    try_values([](Tuning& t, float v) { t.grid_scale = v; }, scales);
    const size_t cap = I.eng->capacity();
    std::vector<size_t> sizes;
    for (size_t div : {2, 4})
        if (sample_paths > cap / div) sizes.push_back(cap / div);
    try_values([](Tuning& t, size_t v) { t.pass_paths = v; }, sizes);
    if (frames) {
        const int shifts[] = {3, 5, 6};
        try_values([](Tuning& t, int v) { t.tile_shift = v; }, shifts);
    }
    const int sorts[] = {1 - cur.ray_sort};
    try_values([](Tuning& t, int v) { t.ray_sort = v; }, sorts);
    set_tuning(cur);
    rep.best = cur;
    rep.best_ms = best_ms;
    return rep;
}

// ------------------------------------------------------------------ serving
namespace {

void wait_until(const Pool& pool, double t) {
    double dt = t - pool.now();
    if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
}

// done: completion times of the request's frames (one for an accumulate job),
// on the same clock as its arrival.
void record(ServeReport& rep, size_t req, double arrival, const std::vector<double>& done) {
    double prev = arrival;
    std::vector<double>& f = rep.frame_ms[req];
    for (double t : done) {
        f.push_back((t - prev) * 1e3);
        prev = t;
    }
    rep.job_ms[req] = (done.back() - arrival) * 1e3;
    rep.frames += (int)done.size();
    rep.jobs += 1;
}

}  // namespace

ServeReport serve(Pool& pool, const std::vector<ServeRequest>& reqs, int mode) {
    ServeReport rep;
    if (reqs.empty()) return rep;
    if (mode != kServePlain && mode != kServeBatched && mode != kServePlainResident)
        throw std::invalid_argument("serve mode must be 0 (plain), 1 (batched) or 2 (plain, resident renderers)");
    rep.frame_ms.resize(reqs.size());
    rep.job_ms.assign(reqs.size(), 0.0);
    for (const auto& rq : reqs) pool.check(rq.job);
    std::vector<size_t> order(reqs.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    const double base = pool.now();
    double last_done = 0;
    if (mode == kServeBatched) {
        const int cap = pool.max_realtime();
        std::deque<size_t> waiting;  // arrived, not yet admitted
        std::map<int, size_t> running;
        int running_rt = 0;
        size_t next = 0;
        while (next < order.size() || !waiting.empty() || !running.empty()) {
            const double t = pool.now() - base;
            while (next < order.size() && reqs[order[next]].arrival_s <= t) waiting.push_back(order[next++]);
            while (!waiting.empty()) {  // admission in arrival order
                const JobDesc& d = reqs[waiting.front()].job;
                if (cap > 0 && d.frames > 0 && running_rt >= cap) break;
                int id = pool.submit(d);
                if (id < 0) break;  // over the HBM budget: admitted when a running job finishes
                running[id] = waiting.front();
                running_rt += d.frames > 0 ? 1 : 0;
                waiting.pop_front();
            }
            if (running.empty()) {
                if (!waiting.empty()) throw std::runtime_error("a request never fits the HBM budget");
                wait_until(pool, base + reqs[order[next]].arrival_s);
                continue;
            }
            pool.step();
            for (auto it = running.begin(); it != running.end();) {
                const JobStatus& st = pool.status(it->first);
                if (!st.done) {
                    ++it;
                    continue;
                }
                const ServeRequest& rq = reqs[it->second];
                std::vector<double> done;
                if (rq.job.frames > 0) {
                    for (double x : st.frame_done_s) done.push_back(x - base);
                } else {
                    done.push_back(st.done_s - base);
                }
                record(rep, it->second, rq.arrival_s, done);
                last_done = std::max(last_done, done.back());
                if (rq.result) {
                    if (rq.job.frames > 0) pool.read_frame(it->first, 0 /* denoised */, rq.result);
                    else pool.read_accum(it->first, 2 /* mean */, rq.result);
                }
                running_rt -= rq.job.frames > 0 ? 1 : 0;
                pool.release(it->first);
                it = running.erase(it);
            }
        }
    } else {
        // One job at a time in arrival order: a fresh single-scene renderer whose
        // queues are sized for its own samples, or (resident) a renderer kept per
        // scene and settings, its history and accumulation reset for each job.
        struct Resident {
            RenderSettings set;
            std::unique_ptr<Renderer> r;
        };
        std::map<int, Resident> resident;
        auto same = [](const RenderSettings& a, const RenderSettings& b) {
            return a.max_bounces == b.max_bounces && a.rr == b.rr && a.rr_start_depth == b.rr_start_depth &&
                   a.seed == b.seed && a.batch_spp == b.batch_spp && a.queue_budget_bytes == b.queue_budget_bytes &&
                   a.flush_survivors == b.flush_survivors && a.env_alias == b.env_alias &&
                   a.use_graphs == b.use_graphs && a.ray_sort == b.ray_sort && a.shadow_any_hit == b.shadow_any_hit &&
                   a.profile == b.profile;
        };
        for (size_t k : order) {
            const ServeRequest& rq = reqs[k];
            wait_until(pool, base + rq.arrival_s);
            const JobDesc& d = rq.job;
            RenderSettings set = d.set;
            // Accumulate jobs run all their samples in one pass (up to 64) unless
            // the job's own settings ask for smaller passes.
            set.batch_spp = d.frames > 0 ? d.spp
                                         : (d.set.batch_spp > 0 ? std::min(d.set.batch_spp, d.spp) : std::min(d.spp, 64));
            std::vector<double> done;
            std::unique_ptr<Renderer> fresh;
            Renderer* r = nullptr;
            if (mode == kServePlainResident) {
                Resident& res = resident[d.scene];
                if (!res.r || !same(res.set, set)) {
                    res.r.reset();
                    res.r.reset(new Renderer(pool.scene_gpu(d.scene), pool.scene_host(d.scene), set));
                    res.set = set;
                }
                r = res.r.get();
                r->reset_history();
                r->reset();
            } else {
                fresh.reset(new Renderer(pool.scene_gpu(d.scene), pool.scene_host(d.scene), set));
                r = fresh.get();
            }
            if (d.frames > 0) {
                for (int f = 0; f < d.frames; ++f) {
                    r->set_camera(d.cams[f]);
                    r->render_frame(d.spp, d.svgf);
                    r->synchronize();
                    done.push_back(pool.now() - base);
                }
                if (rq.result) r->read_frame(0 /* denoised */, rq.result);
            } else {
                if (!d.cams.empty()) r->set_camera(d.cams[0]);
                r->render(d.spp);
                r->synchronize();
                done.push_back(pool.now() - base);
                if (rq.result) r->read_accum(2 /* mean */, rq.result);
            }
            fresh.reset();
            record(rep, k, rq.arrival_s, done);
            last_done = std::max(last_done, done.back());
        }
    }
    rep.makespan_s = last_done - reqs[order.front()].arrival_s;
    return rep;
}

}  // namespace rt
