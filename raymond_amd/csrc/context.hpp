// The library's internal state: device buffers, scenes and contexts, shared
// by the C-ABI layer (api.hip) and the job server (server.hip).
#pragma once
#include <raymond.h>

#include <string>
#include <vector>

#include "kernels.hpp"

namespace rmi {

// Sets the calling thread's rm_last_error() message and returns code.
int fail(int code, const std::string& message);
int hip_fail(hipError_t e, const char* what);

#define RM_HIP(call, what)                                     \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::rmi::hip_fail(e_, what); \
    } while (0)

// The streams a context's buffers are used on, for the block cache.
struct Owner {
    int device = 0;
    hipStream_t streams[3] = {nullptr, nullptr, nullptr};
};

// The device's block cache. hipFree waits for every stream of the device, so
// a context that gave back a buffer (to grow it, or to trim a job server's
// slot) would drain the whole GPU, all other contexts' frames with it. A
// context's buffers are given back to this cache instead, with events
// recorded on the context's streams; cache_take hands a block on only once
// those events have passed, so no work still using it can see it reused.
// Blocks are reused for requests of at least half their size. The cache holds
// at most BLOCK_CACHE_BYTES a device; past that, and when an allocation fails,
// it frees its oldest blocks (after their events).
constexpr uint64_t BLOCK_CACHE_BYTES = 16ull << 30;
hipError_t cache_take(int device, size_t need, void** ptr, size_t* bytes);
void cache_put(const Owner& owner, void* ptr, size_t bytes);
uint64_t cache_bytes(int device);
// Called by a context that is closing, once its streams are idle: drops the
// events its released blocks carry (they have passed), before the streams go.
void cache_forget(const Owner& owner);
// Frees every cached block of the device (waiting for their events).
void cache_flush(int device);

// A growable device allocation. With an owner (a context's buffer), memory
// comes from and goes back to the device's block cache; without one (a
// scene's buffers, read by every context), straight from hipMalloc and to
// hipFree.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    const Owner* owner = nullptr;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        release();
        hipError_t e = owner ? cache_take(owner->device, need, &ptr, &bytes) : hipMalloc(&ptr, need);
        if (e == hipSuccess) {
            if (!owner) bytes = need;
        } else {
            // The failure is returned here; left as the thread's last error,
            // it would be reported again by the next launch's check.
            ptr = nullptr;
            bytes = 0;
            (void)hipGetLastError();
        }
        return e;
    }
    void release() {
        if (ptr) {
            if (owner) cache_put(*owner, ptr, bytes);
            else (void)hipFree(ptr);
        }
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct StageTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[RM_STAGE_COUNT];
    std::vector<hipEvent_t> pool;
    double ms[RM_STAGE_COUNT] = {};
    uint32_t launches[RM_STAGE_COUNT] = {};
};

}  // namespace rmi

struct rm_scene {
    rmi::DevBuf prims, nodes, mats, attr, lights, textures, nmaps;
    rmi::DevBuf qnodes;   // the 8-bit child-quantized tree; empty when the tree can't be quantized
    rmi::DevBuf wnodes;   // the 4-wide quantized tree; empty when the tree can't be folded
    // The delta cut (rm::DevScene::delta_boxes): delta_count box pairs, or
    // rm::NO_DELTA_CUT when the nodes gave none (see delta_cut in api.hip).
    rmi::DevBuf delta;
    uint32_t delta_count = rm::NO_DELTA_CUT;
    uint32_t primitive_slots = 0, primitive_count = 0, node_count = 0, material_count = 0;
    uint32_t tex_w = 0, tex_h = 0, tex_layers = 0, nmap_w = 0, nmap_h = 0, nmap_layers = 0;
    bool dispersive = false;
    std::vector<uint32_t> lights_host;   // the lights table, for checking parameters against
    std::vector<uint8_t> kinds_host;     // each primitive slot's kind, for checking emitters against
    void release() {
        for (rmi::DevBuf* b : {&prims, &nodes, &mats, &attr, &lights, &textures, &nmaps, &qnodes, &wnodes, &delta})
            b->release();
    }
};

static_assert(RM_OPTION_SCHEDULE == 0, "the schedule is the first option");

struct rm_context {
    int device = 0;
    rmi::Owner owner;   // the device and streams its buffers are used on
    hipStream_t stream = nullptr;
    // Light tracing and photon gather run on `side`, alongside the path pass
    // on `stream`; join orders a frame's resolve after them.
    hipStream_t side = nullptr;
    hipEvent_t join = nullptr;
    // The denoiser's à-trous passes and compose run on `post` once prepare
    // has read the frame's accumulation, so the next frame's path pass can
    // start meanwhile; `prepared_ev` and `denoised_ev` order the two.
    hipStream_t post = nullptr;
    hipEvent_t prepared_ev = nullptr, denoised_ev = nullptr;
    hipEvent_t batch_ev = nullptr;   // orders a batched path pass (rm_frame_batch) with this context's stream
    hipEvent_t side_ev = nullptr;    // orders batched light tracing (rm_frame_batch) with this context's side stream
    hipEvent_t tune_ev[2] = {nullptr, nullptr};   // rm_tune's timing events on the path stream
    bool denoise_pending = false;
    uint32_t width = 0, height = 0, capacity = 0;
    rmi::DevBuf accum, galb, gnorm, light_a, light_b, denoised, guides, albedo_frame;
    // Light tracing's splats and the photon gather's output, a pair per frame
    // parity: the side stream fills one pair while the path stream's resolve
    // reads the other, so a frame's light tracing waits only for the resolve
    // that last read its pair (resolved_ev), not for the path pass before it.
    rmi::DevBuf splats[2], gathered[2];
    hipEvent_t resolved_ev[2] = {nullptr, nullptr};
    // The denoiser's end for the last frame that used each pair (recorded
    // when that frame denoised): a frame's light tracing also waits for it.
    hipEvent_t denoised_pair_ev[2] = {nullptr, nullptr};
    bool denoised_pair_live[2] = {false, false};
    uint32_t pair = 0;   // the pair the next frame with light tracing uses
    // The photon map (rm::PhotonMap): cell counts, cell starts, and the
    // photons' deposits with their sorted copy (96 bytes a photon slot).
    rmi::DevBuf photon_counts, photon_starts, photons, emitters, spill, side_spill;
    rmi::DevBuf gather_queue;   // the photon gather's queue (rm::GatherQueue): a slot a pixel, then its length
    std::vector<uint8_t> emitters_host;   // the bytes now in `emitters`
    rmi::DevBuf env, env_table;
    uint32_t env_w = 1, env_h = 1, table_w = 1, table_h = 0;
    rmi::DevBuf rays, hits, counters, work, staging, display;
    // Compensated sums (RM_PRECISION_COMPENSATED): comp holds 3 float4 of
    // rounding error a pixel for accum, galb and gnorm; comp_live says the
    // last frame kept it up to date; folded holds sum + error for readers.
    rmi::DevBuf comp, folded;
    rmi::DevBuf wavefront;   // RM_SCHEDULE_WAVEFRONT's path state and queues
    rmi::DevBuf split;       // RM_SCHEDULE_SPLIT's queue of paths and its counters
    bool comp_live = false;
    rm::Params last_params = {};   // the last frame's, for rm_display
    bool rendered = false;
    int compute_units = 0;
    // This device's copy of the static raycast scratch (rm_raycast_static),
    // looked up by rm_open.
    uint2* static_scratch = nullptr;
    // A context starts on the split schedule (the first option); every other
    // option starts at 0, its default.
    int64_t options[RM_OPTION_COUNT] = {RM_SCHEDULE_SPLIT};
    bool timing = false;
    rmi::StageTimer timer;
    std::vector<rm_scene*> scenes;
};

namespace rmi {
// The per-pixel sums (accum, guide albedo, guide normal) at their best FP32
// value, ordered on `stream`: the running sums themselves, or when the last
// frame kept compensated sums, sum + error folded into ctx->folded.
int current_sums(rm_context* ctx, hipStream_t stream, const float4* out[3]);
// The tree a frame on ctx walks for scene (rm::TREE_*).
int tree_of(const rm_context* ctx, const rm_scene* scene);
// Bytes of device memory the context's own buffers hold (its scenes apart).
uint64_t context_bytes(const rm_context* ctx);
// Waits for the context's streams and frees its per-image and working
// buffers (the next rm_resize and frames allocate them again at their size).
int trim(rm_context* ctx);
}  // namespace rmi
