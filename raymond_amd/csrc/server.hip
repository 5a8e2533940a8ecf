// The job server (rm_server_* in include/raymond.h): a queue of render jobs,
// admitted into a fixed number of slots as slots and memory free up, each
// active job rendering in a context of its own so that the frames of
// different jobs overlap on the device. One host thread schedules: it admits
// jobs, keeps two frames of every active job in flight, notices finished
// frames by event queries, and copies a finished job's image to pinned host
// memory. A job's frames are exactly the frames rm_frame renders on a
// context of its own, so its image does not depend on what else runs.
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "context.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Word offsets in the parameter block (src/shaders.mjs:9-24).
constexpr int P_WIDTH = 40, P_HEIGHT = 41, P_SAMPLE_OFFSET = 42, P_SPP = 43, P_TOTAL = 44, P_DENOISE = 46,
              P_CAUSTICS = 47, P_RADIUS = 48, P_CELL = 49;

struct Job {
    uint64_t id = 0;
    const rm_scene* scene = nullptr;
    uint32_t params[RM_PARAMS_BYTES / 4] = {};
    std::vector<float> emitters;
    std::vector<uint16_t> env_rgba;
    std::vector<float> env_table;
    uint32_t env_w = 0, env_h = 0, table_w = 0, table_h = 0;
    uint32_t width = 0, height = 0, spp = 1, frames_total = 0;
    bool denoise = false, caustics = false;
    float radius0 = 0.0f;
    uint64_t bytes = 0;
    int64_t options[RM_OPTION_COUNT] = {};   // the server's options when the job was submitted
    bool has_deadline = false;
    Clock::time_point deadline;
    // Set by rm_server_cancel (CANCEL_ASKED) or on a missed deadline
    // (CANCEL_LATE): no more frames are queued, and the job ends with
    // RM_ERR_CANCELLED once those in flight are done.
    std::atomic<int> cancel{0};
    // progress, owned by the scheduler thread until `finished`
    uint32_t frames_enqueued = 0, frames_done = 0;
    Clock::time_point t_submit, t_admit, t_first, t_done;
    bool readback = false, finished = false;
    int error = RM_OK;
    std::string message;
    std::vector<float> image;
};

struct Slot {
    rm_context* ctx = nullptr;
    std::shared_ptr<Job> job;
    hipEvent_t frame_ev[2] = {nullptr, nullptr};  // the in-flight frames, oldest at `head`
    uint32_t head = 0, in_flight = 0;
    hipEvent_t readback_ev = nullptr;
    float* pinned = nullptr;
    size_t pinned_floats = 0;
};

constexpr int CANCEL_ASKED = 1, CANCEL_LATE = 2;

std::string cancel_message(int why) {
    return why == CANCEL_LATE ? "it missed its deadline." : "it was cancelled.";
}

// Device memory a job's context needs, by the buffers rm_frame (and rm_tune)
// size for it. options: the job's RM_OPTION_* values.
uint64_t job_bytes(uint32_t w, uint32_t h, bool caustics, const int64_t* options, uint64_t env_bytes, int cus) {
    const uint64_t px = (uint64_t)w * h;
    const bool compensated = (options[RM_OPTION_PRECISION] & RM_PRECISION_COMPENSATED) != 0;
    const int64_t schedule = options[RM_OPTION_SCHEDULE];
    const bool tuned = options[RM_OPTION_AUTOTUNE] > 0;   // rm_tune tries every schedule
    // accum, two guides, two denoiser buffers, denoised, two splat buffers,
    // guides, albedo: 10 float4 a pixel, and the traversal spill.
    uint64_t b = px * (10 * 16 + rm::SPILL_ENTRIES * 8) + env_bytes;
    if (compensated) b += px * 96;   // error terms and their folded sums
    // The compacted schedule's staging records (16×16 tiles rounded up) and
    // the wavefront schedule's path state and queues.
    const uint64_t tiled = (uint64_t)((w + 15) / 16) * ((h + 15) / 16) * rm::BLOCK_THREADS;
    if (tuned || schedule == RM_SCHEDULE_COMPACT || schedule == RM_SCHEDULE_SORTED) b += tiled * 48 + (tiled > px ? tiled - px : 0) * rm::SPILL_ENTRIES * 8;
    if (tuned || schedule == RM_SCHEDULE_WAVEFRONT) b += rm::wavefront_bytes(px);
    if (tuned || schedule == RM_SCHEDULE_SPLIT) {
        // The split schedule's queue, and its tail's spill lanes past the pixels'.
        const uint64_t tail = (uint64_t)rm::split_tail_blocks(cus, 5) * rm::BLOCK_THREADS;
        b += rm::split_bytes(px) + (tail > px ? tail - px : 0) * rm::SPILL_ENTRIES * 8;
    }
    if (caustics) {
        const uint64_t budget = (uint64_t)options[RM_OPTION_LIGHT_PATHS];
        const uint64_t photons = budget > rm::PHOTON_CAPACITY ? budget : rm::PHOTON_CAPACITY;
        const uint64_t lanes = photons < rm::LIGHT_LANES ? photons : rm::LIGHT_LANES;
        b += px * 2 * 16;   // the two gathered-light buffers
        b += px * rm::GATHER_SLOT_VEC4S * 16 + 256;   // the gather queue
        b += (px > lanes ? px : lanes) * rm::SPILL_ENTRIES * 8;
        b += (uint64_t)(rm::PHOTON_TABLE + 1) * 4 + rm::photon_map_bytes(photons);
    }
    return b;
}

}  // namespace

struct rm_server {
    int device = 0;
    uint64_t budget = 0, used = 0;
    int64_t options[RM_OPTION_COUNT] = {RM_SCHEDULE_SPLIT};   // given to jobs submitted from now on
    // The schedule rm_tune picked for each (scene, width, height, samples per
    // frame), with RM_OPTION_AUTOTUNE; used by the scheduler thread only.
    // The options rm_tune set: schedule, persistent blocks, waves, stack depth, block shape, split depth.
    std::map<std::tuple<const rm_scene*, uint32_t, uint32_t, uint32_t>, std::array<int64_t, 6>> tuned;
    rm_context* home = nullptr;
    std::vector<Slot> slots;
    std::mutex m;
    std::condition_variable wake, done;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;

    void run();
    void drop_queued();
    bool admit();
    bool progress(Slot& s, bool enqueue);
    bool enqueue_batched();
    void finish(Slot& s, int error, const std::string& message);
};

namespace {

// Whether frame k of a job runs the denoiser: only its first frame (the
// time-to-first-image a client sees) and its last (the result). The denoiser
// reads nothing but the current accumulation and guides, so skipping it in
// between leaves the final image unchanged.
bool denoises(const Job& j, uint32_t k) { return j.denoise && (k == 0 || k + 1 == j.frames_total); }

// The parameter block of a job's frame k, advanced as the Renderer advances
// it: sample offset and running total, and the photon radius shrinking as
// (k + 1)^(-1/6).
void frame_params(const Job& j, uint32_t k, uint32_t* out) {
    std::memcpy(out, j.params, RM_PARAMS_BYTES);
    out[P_DENOISE] = denoises(j, k) ? 1u : 0u;
    out[P_SAMPLE_OFFSET] = k * j.spp;
    out[P_TOTAL] = (k + 1) * j.spp;
    float r = (float)((double)j.radius0 * std::pow((double)(k + 1), -1.0 / 6.0));
    float cell = (float)(2.0 * (double)r);
    std::memcpy(&out[P_RADIUS], &r, 4);
    std::memcpy(&out[P_CELL], &cell, 4);
}

}  // namespace

void rm_server::finish(Slot& s, int error, const std::string& message) {
    std::shared_ptr<Job> j = s.job;
    {
        std::lock_guard<std::mutex> lock(m);
        j->error = error;
        j->message = message;
        j->t_done = Clock::now();
        j->finished = true;
        used -= j->bytes;
        s.job.reset();
        s.in_flight = 0;
        s.head = 0;
    }
    done.notify_all();
}

// Ends queued jobs that were cancelled or missed their deadline, without a slot.
void rm_server::drop_queued() {
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lock(m);
    bool dropped = false;
    for (auto it = queue.begin(); it != queue.end();) {
        Job& j = **it;
        if (j.cancel == 0 && j.has_deadline && now > j.deadline) j.cancel = CANCEL_LATE;
        if (j.cancel != 0) {
            j.error = RM_ERR_CANCELLED;
            j.message = cancel_message(j.cancel);
            j.t_admit = j.t_done = now;
            j.finished = true;
            it = queue.erase(it);
            dropped = true;
        } else {
            ++it;
        }
    }
    if (dropped) done.notify_all();
}

// Moves queued jobs into free slots while their memory fits. Returns whether
// any job was admitted.
bool rm_server::admit() {
    bool any = false;
    for (Slot& s : slots) {
        if (s.job) continue;
        std::shared_ptr<Job> j;
        {
            std::lock_guard<std::mutex> lock(m);
            if (queue.empty()) return any;
            bool busy = false;
            for (Slot& o : slots) busy = busy || (bool)o.job;
            // A job larger than the whole budget still runs, alone. Blocks
            // the slots gave back to the device's cache count against the
            // budget too: when they would push it over, they are freed.
            if (used + queue.front()->bytes + rmi::cache_bytes(device) > budget) rmi::cache_flush(device);
            if (busy && used + queue.front()->bytes > budget) return any;
            j = queue.front();
            queue.pop_front();
            used += j->bytes;
            s.job = j;
        }
        j->t_admit = Clock::now();
        any = true;
        int r = RM_OK;
        // A slot keeps the buffers of the largest job it ran; charged only
        // this job's size, it could hold more than the budget allows. So a
        // slot holding well more than this job needs gives its buffers back
        // first (a slot about the right size keeps them: no reallocation).
        if (rmi::context_bytes(s.ctx) > j->bytes + j->bytes / 4 + (64ull << 20)) r = rmi::trim(s.ctx);
        for (int k = 0; k < RM_OPTION_COUNT && r == RM_OK; ++k) r = rm_set_option(s.ctx, k, j->options[k]);
        if (r == RM_OK) r = rm_resize(s.ctx, j->width, j->height);
        if (r == RM_OK) {
            if (!j->env_rgba.empty())
                r = rm_environment(s.ctx, j->env_rgba.data(), j->env_w, j->env_h, j->env_table.data(), j->table_w,
                                   j->table_h);
            else {
                const uint16_t black[4] = {0, 0, 0, 0x3c00};
                const float table[1] = {0.0f};
                r = rm_environment(s.ctx, black, 1, 1, table, 1, 0);
            }
        }
        if (r == RM_OK && j->options[RM_OPTION_AUTOTUNE] > 0) {
            // Every candidate gives the same image, so the choice changes speed only.
            const auto key = std::make_tuple(j->scene, j->width, j->height, j->spp);
            auto it = tuned.find(key);
            if (it == tuned.end()) {
                uint32_t p[RM_PARAMS_BYTES / 4];
                frame_params(*j, 0, p);
                rm_tune_result res;
                r = rm_tune(s.ctx, j->scene, p, j->caustics ? j->emitters.data() : nullptr,
                            j->caustics ? (uint32_t)(j->emitters.size() / 4) : 0u,
                            (uint32_t)j->options[RM_OPTION_AUTOTUNE], &res);
                if (r == RM_OK)
                    it = tuned.emplace(key, std::array<int64_t, 6>{res.schedule[res.best], res.persistent_blocks[res.best],
                                                                   res.waves[res.best], res.stack_depth[res.best],
                                                                   res.block_shape[res.best], res.split_depth[res.best]})
                             .first;
            }
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_SCHEDULE, it->second[0]);
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_PERSISTENT_BLOCKS, it->second[1]);
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_WAVES, it->second[2]);
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_STACK_DEPTH, it->second[3]);
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_BLOCK_SHAPE, it->second[4]);
            if (r == RM_OK) r = rm_set_option(s.ctx, RM_OPTION_SPLIT_DEPTH, it->second[5]);
        }
        if (r == RM_OK) {
            // The job is charged what its slot now holds, if that is more
            // than the estimate (its frames only grow light tracing's buffers,
            // which the estimate counts).
            const uint64_t held = rmi::context_bytes(s.ctx);
            std::lock_guard<std::mutex> lock(m);
            if (held > j->bytes) {
                used += held - j->bytes;
                j->bytes = held;
            }
        }
        if (r != RM_OK) finish(s, r, rm_last_error());
    }
    return any;
}

// Advances one slot: retires finished frames, enqueues more, reads back a
// finished job. Returns whether anything changed.
bool rm_server::progress(Slot& s, bool enqueue) {
    if (!s.job) return false;
    Job& j = *s.job;
    bool changed = false;
    // A deadline is missed only while frames remain to be queued; once the
    // last is queued the job finishes with its image.
    if (j.cancel == 0 && j.has_deadline && j.frames_enqueued < j.frames_total && Clock::now() > j.deadline)
        j.cancel = CANCEL_LATE;
    const int cancelled = j.frames_enqueued < j.frames_total ? j.cancel.load() : 0;
    if (cancelled) enqueue = false;
    while (s.in_flight > 0) {
        hipError_t q = hipEventQuery(s.frame_ev[s.head]);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
            finish(s, RM_ERR_DEVICE, std::string("a frame failed: ") + hipGetErrorString(q));
            return true;
        }
        if (j.frames_done == 0) j.t_first = Clock::now();
        j.frames_done++;
        s.head ^= 1u;
        s.in_flight--;
        changed = true;
    }
    while (enqueue && s.in_flight < 2 && j.frames_enqueued < j.frames_total) {
        uint32_t p[RM_PARAMS_BYTES / 4];
        frame_params(j, j.frames_enqueued, p);
        int r = rm_frame(s.ctx, j.scene, p, j.caustics ? j.emitters.data() : nullptr,
                         (uint32_t)(j.emitters.size() / 4));
        if (r != RM_OK) {
            finish(s, r, rm_last_error());
            return true;
        }
        // The frame is complete when its last stream is: the denoiser's when
        // it runs (it follows the path pass, resolve and prepare), else the path stream's.
        hipStream_t last = denoises(j, j.frames_enqueued) ? s.ctx->post : s.ctx->stream;
        hipError_t e = hipEventRecord(s.frame_ev[(s.head + s.in_flight) & 1u], last);
        if (e != hipSuccess) {
            finish(s, RM_ERR_DEVICE, std::string("recording a frame event: ") + hipGetErrorString(e));
            return true;
        }
        j.frames_enqueued++;
        s.in_flight++;
        changed = true;
    }
    if (cancelled && s.in_flight == 0) {
        finish(s, RM_ERR_CANCELLED, cancel_message(cancelled));
        return true;
    }
    if (j.frames_done == j.frames_total && !j.readback) {
        const size_t floats = (size_t)j.width * j.height * 4;
        if (s.pinned_floats < floats) {
            if (s.pinned) (void)hipHostFree(s.pinned);
            s.pinned = nullptr;
            s.pinned_floats = 0;
            hipError_t e = hipHostMalloc((void**)&s.pinned, floats * 4, hipHostMallocDefault);
            if (e != hipSuccess) {
                s.pinned = nullptr;
                (void)hipGetLastError();
                finish(s, RM_ERR_MEMORY, std::string("allocating pinned memory: ") + hipGetErrorString(e));
                return true;
            }
            s.pinned_floats = floats;
        }
        hipStream_t last = j.denoise ? s.ctx->post : s.ctx->stream;
        const void* src = s.ctx->denoised.ptr;
        if (!j.denoise) {
            const float4* sums[3];
            if (rmi::current_sums(s.ctx, last, sums) != RM_OK) {
                finish(s, RM_ERR_DEVICE, std::string("reading back an image: ") + rm_last_error());
                return true;
            }
            src = sums[0];
        }
        hipError_t e = hipMemcpyAsync(s.pinned, src, floats * 4, hipMemcpyDeviceToHost, last);
        if (e == hipSuccess) e = hipEventRecord(s.readback_ev, last);
        if (e != hipSuccess) {
            finish(s, RM_ERR_DEVICE, std::string("reading back an image: ") + hipGetErrorString(e));
            return true;
        }
        j.readback = true;
        changed = true;
    }
    if (j.readback) {
        hipError_t q = hipEventQuery(s.readback_ev);
        if (q == hipSuccess) {
            j.image.assign(s.pinned, s.pinned + (size_t)j.width * j.height * 4);
            finish(s, RM_OK, "");
            changed = true;
        } else if (q != hipErrorNotReady) {
            finish(s, RM_ERR_DEVICE, std::string("reading back an image: ") + hipGetErrorString(q));
            changed = true;
        }
    }
    return changed;
}

// With RM_OPTION_BATCH: the next frame of every slot with room, in batches
// of jobs that walk the same tree, one path-pass launch a batch.
bool rm_server::enqueue_batched() {
    bool changed = false;
    for (;;) {
        std::vector<Slot*> ready;
        for (Slot& s : slots)
            if (s.job && s.in_flight < 2 && s.job->frames_enqueued < s.job->frames_total && s.job->cancel == 0)
                ready.push_back(&s);
        if (ready.empty()) return changed;
        // A batch walks one kind of tree, and runs the per-pixel pass or,
        // when all its jobs are on the wavefront schedule, the wavefront one.
        auto wavefront = [](const Slot* s) { return s->ctx->options[RM_OPTION_SCHEDULE] == RM_SCHEDULE_WAVEFRONT; };
        const int tree = rmi::tree_of(ready[0]->ctx, ready[0]->job->scene);
        const bool wf = wavefront(ready[0]);
        std::vector<Slot*> group;
        for (Slot* s : ready)
            if (group.size() < rm::MAX_BATCH && rmi::tree_of(s->ctx, s->job->scene) == tree && wavefront(s) == wf)
                group.push_back(s);
        const uint32_t n = (uint32_t)group.size();
        rm_context* ctxs[rm::MAX_BATCH];
        const rm_scene* scenes[rm::MAX_BATCH];
        uint32_t params[rm::MAX_BATCH][RM_PARAMS_BYTES / 4];
        const void* pp[rm::MAX_BATCH];
        const float* em[rm::MAX_BATCH];
        uint32_t ev[rm::MAX_BATCH];
        for (uint32_t i = 0; i < n; ++i) {
            const Job& j = *group[i]->job;
            ctxs[i] = group[i]->ctx;
            scenes[i] = j.scene;
            frame_params(j, j.frames_enqueued, params[i]);
            pp[i] = params[i];
            em[i] = j.caustics ? j.emitters.data() : nullptr;
            ev[i] = j.caustics ? (uint32_t)(j.emitters.size() / 4) : 0u;
        }
        int r = rm_frame_batch(ctxs, scenes, pp, em, ev, n);
        changed = true;
        if (r != RM_OK) {
            const std::string message = rm_last_error();
            for (Slot* s : group) finish(*s, r, message);
            continue;
        }
        for (Slot* s : group) {
            Job& j = *s->job;
            hipStream_t last = denoises(j, j.frames_enqueued) ? s->ctx->post : s->ctx->stream;
            hipError_t e = hipEventRecord(s->frame_ev[(s->head + s->in_flight) & 1u], last);
            if (e != hipSuccess) {
                finish(*s, RM_ERR_DEVICE, std::string("recording a frame event: ") + hipGetErrorString(e));
                continue;
            }
            j.frames_enqueued++;
            s->in_flight++;
        }
    }
}

void rm_server::run() {
    (void)hipSetDevice(device);
    for (;;) {
        drop_queued();
        bool changed = admit();
        bool active = false;
        bool batched;
        {
            std::lock_guard<std::mutex> lock(m);
            batched = options[RM_OPTION_BATCH] != 0;
        }
        for (Slot& s : slots) {
            changed = progress(s, !batched) || changed;
            active = active || (bool)s.job;
        }
        if (batched) changed = enqueue_batched() || changed;
        if (!active) {
            std::unique_lock<std::mutex> lock(m);
            if (stop && queue.empty()) return;
            if (queue.empty()) wake.wait(lock, [&] { return stop || !queue.empty(); });
            continue;
        }
        // Frames are in flight: look again shortly rather than spin.
        if (!changed) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

extern "C" {

int rm_server_open(int device, uint32_t slots, uint64_t memory_budget, rm_server** out) {
    if (!out) return rmi::fail(RM_ERR_INVALID, "rm_server_open needs somewhere to put the server.");
    *out = nullptr;
    if (slots == 0 || slots > 64) return rmi::fail(RM_ERR_INVALID, "A server has 1 to 64 slots.");
    std::unique_ptr<rm_server> sv(new rm_server());
    sv->device = device;
    sv->budget = memory_budget;
    int r = rm_open(device, &sv->home);
    if (r != RM_OK) return r;
    sv->slots.resize(slots);
    for (Slot& s : sv->slots) {
        r = rm_open(device, &s.ctx);
        hipError_t e = hipSuccess;
        if (r == RM_OK) e = hipEventCreateWithFlags(&s.frame_ev[0], hipEventDisableTiming);
        if (r == RM_OK && e == hipSuccess) e = hipEventCreateWithFlags(&s.frame_ev[1], hipEventDisableTiming);
        if (r == RM_OK && e == hipSuccess) e = hipEventCreateWithFlags(&s.readback_ev, hipEventDisableTiming);
        if (r != RM_OK || e != hipSuccess) {
            int code = r != RM_OK ? r : rmi::hip_fail(e, "creating the server's events");
            std::string message = rm_last_error();
            for (Slot& o : sv->slots) {
                for (hipEvent_t ev : {o.frame_ev[0], o.frame_ev[1], o.readback_ev})
                    if (ev) (void)hipEventDestroy(ev);
                if (o.ctx) rm_close(o.ctx);
            }
            rm_close(sv->home);
            return rmi::fail(code, message);
        }
    }
    sv->worker = std::thread([p = sv.get()] { p->run(); });
    *out = sv.release();
    return RM_OK;
}

rm_context* rm_server_context(rm_server* server) { return server ? server->home : nullptr; }

int rm_server_submit(rm_server* sv, const rm_job* job, uint64_t* id) {
    if (!sv || !job || !id) return rmi::fail(RM_ERR_INVALID, "rm_server_submit needs a server, a job and an id.");
    if (!job->scene || !job->params) return rmi::fail(RM_ERR_INVALID, "A job needs a scene and parameters.");
    auto j = std::make_shared<Job>();
    j->scene = job->scene;
    std::memcpy(j->params, job->params, RM_PARAMS_BYTES);
    j->width = j->params[P_WIDTH];
    j->height = j->params[P_HEIGHT];
    j->spp = j->params[P_SPP];
    j->denoise = j->params[P_DENOISE] != 0u;
    j->caustics = j->params[P_CAUSTICS] != 0u;
    std::memcpy(&j->radius0, &j->params[P_RADIUS], 4);
    if (j->width == 0 || j->height == 0 || j->spp == 0)
        return rmi::fail(RM_ERR_INVALID, "A job's parameters need a size and at least one sample per frame.");
    if (job->target_samples == 0) return rmi::fail(RM_ERR_INVALID, "A job needs a target of at least one sample.");
    if (!(job->deadline_ms >= 0.0))
        return rmi::fail(RM_ERR_INVALID, "A job's deadline is a number of milliseconds, 0 for none.");
    j->frames_total = (job->target_samples + j->spp - 1) / j->spp;
    if (j->caustics) {
        if (!job->emitters || job->emitter_vec4s < 2)
            return rmi::fail(RM_ERR_INVALID, "A job with light tracing needs its emitters.");
        j->emitters.assign(job->emitters, job->emitters + (size_t)job->emitter_vec4s * 4);
    }
    if (job->env_rgba) {
        if (!job->env_table || job->env_width == 0 || job->env_height == 0 || job->table_width == 0)
            return rmi::fail(RM_ERR_INVALID, "A job's environment needs texels and an importance table.");
        j->env_rgba.assign(job->env_rgba, job->env_rgba + (size_t)job->env_width * job->env_height * 4);
        j->env_table.assign(job->env_table, job->env_table + (size_t)job->table_width * (job->table_height + 1));
        j->env_w = job->env_width;
        j->env_h = job->env_height;
        j->table_w = job->table_width;
        j->table_h = job->table_height;
    }
    {
        std::lock_guard<std::mutex> lock(sv->m);
        std::memcpy(j->options, sv->options, sizeof j->options);
    }
    const uint64_t env_bytes = j->env_rgba.empty()
                                   ? 0
                                   : (uint64_t)j->env_w * j->env_h * 8 + (uint64_t)j->table_w * (j->table_h + 1) * 4;
    j->bytes = job_bytes(j->width, j->height, j->caustics, j->options, env_bytes, sv->home->compute_units);
    // Scene uploads on the home context must have landed before any slot reads them.
    RM_HIP(hipSetDevice(sv->device), "selecting the device");
    RM_HIP(hipStreamSynchronize(sv->home->stream), "waiting for scene uploads");
    j->t_submit = Clock::now();
    if (job->deadline_ms > 0.0) {
        j->has_deadline = true;
        j->deadline = j->t_submit + std::chrono::duration_cast<Clock::duration>(
                                        std::chrono::duration<double, std::milli>(job->deadline_ms));
    }
    {
        std::lock_guard<std::mutex> lock(sv->m);
        if (sv->stop) return rmi::fail(RM_ERR_STATE, "The server is closing.");
        j->id = sv->next_id++;
        sv->jobs[j->id] = j;
        sv->queue.push_back(j);
    }
    sv->wake.notify_all();
    *id = j->id;
    return RM_OK;
}

int rm_server_wait(rm_server* sv, uint64_t id, float* dst, size_t floats, rm_job_stats* stats) {
    if (!sv) return rmi::fail(RM_ERR_INVALID, "rm_server_wait needs a server.");
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lock(sv->m);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return rmi::fail(RM_ERR_INVALID, "No job " + std::to_string(id) + " is known.");
        j = it->second;
        sv->done.wait(lock, [&] { return j->finished; });
        sv->jobs.erase(it);
    }
    if (stats) {
        stats->queued_ms = ms_between(j->t_submit, j->t_admit);   // a job dropped from the queue: until then
        stats->first_frame_ms = j->frames_done ? ms_between(j->t_submit, j->t_first) : 0.0;
        stats->done_ms = ms_between(j->t_submit, j->t_done);
        stats->samples = j->frames_done * j->spp;
        stats->frames = j->frames_done;
    }
    if (j->error != RM_OK) return rmi::fail(j->error, "Job " + std::to_string(id) + " failed: " + j->message);
    if (dst) {
        if (floats > j->image.size()) return rmi::fail(RM_ERR_INVALID, "rm_server_wait asked for more than the image holds.");
        std::memcpy(dst, j->image.data(), floats * 4);
    }
    return RM_OK;
}

int rm_server_cancel(rm_server* sv, uint64_t id) {
    if (!sv) return rmi::fail(RM_ERR_INVALID, "rm_server_cancel needs a server.");
    {
        std::lock_guard<std::mutex> lock(sv->m);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return rmi::fail(RM_ERR_INVALID, "No job " + std::to_string(id) + " is known.");
        int expected = 0;
        it->second->cancel.compare_exchange_strong(expected, CANCEL_ASKED);
    }
    sv->wake.notify_all();
    return RM_OK;
}

int rm_server_set_option(rm_server* sv, int option, int64_t value) {
    if (!sv) return rmi::fail(RM_ERR_INVALID, "rm_server_set_option needs a server.");
    // The home context checks the value exactly as rm_set_option would.
    int r = rm_set_option(sv->home, option, value);
    if (r != RM_OK) return r;
    std::lock_guard<std::mutex> lock(sv->m);
    sv->options[option] = value;
    return RM_OK;
}

int rm_server_close(rm_server* sv) {
    if (!sv) return rmi::fail(RM_ERR_INVALID, "rm_server_close needs a server.");
    {
        std::lock_guard<std::mutex> lock(sv->m);
        sv->stop = true;
    }
    sv->wake.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (Slot& s : sv->slots) {
        for (hipEvent_t ev : {s.frame_ev[0], s.frame_ev[1], s.readback_ev})
            if (ev) (void)hipEventDestroy(ev);
        if (s.pinned) (void)hipHostFree(s.pinned);
        if (s.ctx) rm_close(s.ctx);
    }
    rm_close(sv->home);
    delete sv;
    return RM_OK;
}

}  // extern "C"
