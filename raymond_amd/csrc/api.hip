// The C ABI (include/raymond.h) over the kernels: device memory, the frame's
// dispatch order (src/Renderer.mjs:441-489), readback and stage timing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"

// The traversal spill of rm_raycast_static, which allocates nothing: stack
// entries beyond the raycast kernel's LDS depth for up to
// RM_RAYCAST_STATIC_MAX_RAYS lanes, laid out [entry][lane] as the kernel
// indexes any spill ((STACK_LIMIT - TRACE_LDS_DEPTH) × 65,536 × 8 B = 28 MiB).
// Each device holds its own copy, at its own address, until the process ends.
// In host code `scratch` only names the symbol: kernels get the device
// address that rm_open looks up, never the name itself.
constexpr size_t STATIC_SCRATCH_ENTRIES =
    (size_t)(rm::STACK_LIMIT - rm::TRACE_LDS_DEPTH) * RM_RAYCAST_STATIC_MAX_RAYS;
static __device__ __align__(256) uint2 scratch[STATIC_SCRATCH_ENTRIES];

namespace {
// Each device's address of `scratch`, looked up once with that device current.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
uint2* g_scratch_ptr[MAX_DEVICES] = {};

// The scratch's address on the current device `device`.
hipError_t static_scratch_of(int device, uint2** out) {
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<uint2*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

namespace rmi {

thread_local std::string g_error;

int fail(int code, const std::string& message) {
    g_error = message;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    int code = (e == hipErrorOutOfMemory) ? RM_ERR_MEMORY : RM_ERR_DEVICE;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
// One device's block cache: released blocks, oldest first, each with the
// events that pass once the work issued before its release is done.
struct Block {
    void* ptr;
    size_t bytes;
    hipEvent_t ready[3];
    const Owner* owner;   // whose streams the events are on; null once they have passed
};
struct BlockCache {
    std::mutex lock;
    std::vector<Block> blocks;
    uint64_t bytes = 0;
};
BlockCache g_cache[MAX_DEVICES];

bool block_ready(const Block& b) {
    for (hipEvent_t e : b.ready)
        if (e && hipEventQuery(e) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    return true;
}

// Waits for a block's events, destroys them and frees its memory.
void block_free(const Block& b) {
    for (hipEvent_t e : b.ready)
        if (e) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    (void)hipFree(b.ptr);
}

// Frees the oldest blocks until the cache holds at most `keep` bytes; the
// caller holds the lock.
void shrink_to(BlockCache& cache, uint64_t keep) {
    size_t n = 0;
    while (n < cache.blocks.size() && cache.bytes > keep) {
        block_free(cache.blocks[n]);
        cache.bytes -= cache.blocks[n].bytes;
        ++n;
    }
    cache.blocks.erase(cache.blocks.begin(), cache.blocks.begin() + (std::ptrdiff_t)n);
    (void)hipGetLastError();
}
}  // namespace

hipError_t cache_take(int device, size_t need, void** ptr, size_t* bytes) {
    *ptr = nullptr;
    *bytes = 0;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    BlockCache& cache = g_cache[device];
    {
        std::lock_guard<std::mutex> guard(cache.lock);
        // The smallest ready block of at least `need` and at most twice it.
        size_t best = cache.blocks.size();
        for (size_t i = 0; i < cache.blocks.size(); ++i) {
            const Block& b = cache.blocks[i];
            if (b.bytes < need || b.bytes / 2 > need) continue;
            if (best < cache.blocks.size() && cache.blocks[best].bytes <= b.bytes) continue;
            if (block_ready(b)) best = i;
        }
        if (best < cache.blocks.size()) {
            const Block b = cache.blocks[best];
            for (hipEvent_t e : b.ready)
                if (e) (void)hipEventDestroy(e);
            cache.blocks.erase(cache.blocks.begin() + (std::ptrdiff_t)best);
            cache.bytes -= b.bytes;
            *ptr = b.ptr;
            *bytes = b.bytes;
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(ptr, need);
    if (e == hipErrorOutOfMemory) {
        // The cache's blocks may be what the device lacks: free them all, then try again.
        (void)hipGetLastError();
        cache_flush(device);
        e = hipMalloc(ptr, need);
    }
    if (e != hipSuccess) {
        *ptr = nullptr;
        return e;
    }
    *bytes = need;
    return hipSuccess;
}

void cache_put(const Owner& owner, void* ptr, size_t bytes) {
    if (owner.device < 0 || owner.device >= MAX_DEVICES) return;
    Block b{ptr, bytes, {nullptr, nullptr, nullptr}, &owner};
    bool recorded = true;
    for (int k = 0; k < 3; ++k) {
        if (!owner.streams[k]) continue;
        if (hipEventCreateWithFlags(&b.ready[k], hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(b.ready[k], owner.streams[k]) != hipSuccess) {
            recorded = false;
            break;
        }
    }
    if (!recorded) {
        // Without its events the block can't be handed on safely: free it
        // (hipFree waits for the device).
        (void)hipGetLastError();
        for (hipEvent_t e : b.ready)
            if (e) (void)hipEventDestroy(e);
        (void)hipFree(ptr);
        return;
    }
    BlockCache& cache = g_cache[owner.device];
    std::lock_guard<std::mutex> guard(cache.lock);
    cache.blocks.push_back(b);
    cache.bytes += bytes;
    if (cache.bytes > BLOCK_CACHE_BYTES) shrink_to(cache, BLOCK_CACHE_BYTES);
}

void cache_forget(const Owner& owner) {
    if (owner.device < 0 || owner.device >= MAX_DEVICES) return;
    BlockCache& cache = g_cache[owner.device];
    std::lock_guard<std::mutex> guard(cache.lock);
    for (Block& b : cache.blocks) {
        if (b.owner != &owner) continue;
        for (hipEvent_t& e : b.ready)
            if (e) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        b.owner = nullptr;
    }
}

uint64_t cache_bytes(int device) {
    if (device < 0 || device >= MAX_DEVICES) return 0;
    std::lock_guard<std::mutex> guard(g_cache[device].lock);
    return g_cache[device].bytes;
}

void cache_flush(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> guard(g_cache[device].lock);
    shrink_to(g_cache[device], 0);
}

}  // namespace rmi

using rmi::DevBuf;
using rmi::fail;
using rmi::g_error;
using rmi::hip_fail;

namespace {

hipEvent_t timer_event(rm_context* c) {
    if (!c->timer.pool.empty()) {
        hipEvent_t e = c->timer.pool.back();
        c->timer.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one launch with events on the context's stream when timing.
struct Timed {
    rm_context* c;
    int stage;
    hipStream_t stream;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(rm_context* ctx, int s, hipStream_t on = nullptr) : c(ctx), stage(s), stream(on ? on : ctx->stream) {
        if (c->timing) {
            a = timer_event(c);
            b = timer_event(c);
            if (a && b) (void)hipEventRecord(a, stream);
        }
    }
    ~Timed() {
        if (c->timing && a && b) {
            (void)hipEventRecord(b, stream);
            c->timer.pending[stage].push_back({a, b});
        }
    }
};

int collect_times(rm_context* c) {
    RM_HIP(hipStreamSynchronize(c->stream), "waiting for the stream");
    RM_HIP(hipStreamSynchronize(c->side), "waiting for the side stream");
    RM_HIP(hipStreamSynchronize(c->post), "waiting for the denoiser's stream");
    for (int s = 0; s < RM_STAGE_COUNT; ++s) {
        for (auto& p : c->timer.pending[s]) {
            float ms = 0.0f;
            RM_HIP(hipEventElapsedTime(&ms, p.first, p.second), "reading a stage timer");
            c->timer.ms[s] += ms;
            c->timer.launches[s] += 1;
            c->timer.pool.push_back(p.first);
            c->timer.pool.push_back(p.second);
        }
        c->timer.pending[s].clear();
    }
    return RM_OK;
}

rm::DevScene dev_scene(const rm_context* c, const rm_scene* s) {
    rm::DevScene S;
    S.prims = s->prims.as<const float4>();
    S.nodes = s->nodes.as<const float4>();
    S.qnodes = s->qnodes.as<const float4>();
    S.wnodes = s->wnodes.as<const float4>();
    S.mats = s->mats.as<const float4>();
    S.attr = s->attr.as<const float4>();
    S.lights = s->lights.as<const uint32_t>();
    S.env = c->env.as<const ushort4>();
    S.env_w = c->env_w;
    S.env_h = c->env_h;
    S.env_table = c->env_table.as<const float>();
    S.table_w = c->table_w;
    S.table_h = c->table_h;
    S.textures = s->textures.as<const uchar4>();
    S.tex_w = s->tex_w;
    S.tex_h = s->tex_h;
    S.tex_layers = s->tex_layers;
    S.nmaps = s->nmaps.as<const uchar4>();
    S.nmap_w = s->nmap_w;
    S.nmap_h = s->nmap_h;
    S.nmap_layers = s->nmap_layers;
    S.delta_boxes = s->delta.as<const float4>();
    S.delta_count = s->delta_count;
    return S;
}

}  // namespace

int rmi::current_sums(rm_context* c, hipStream_t stream, const float4* out[3]) {
    if (!c->comp_live) {
        out[0] = c->accum.as<const float4>();
        out[1] = c->galb.as<const float4>();
        out[2] = c->gnorm.as<const float4>();
        return RM_OK;
    }
    const size_t px = (size_t)c->width * c->height;
    RM_HIP(c->folded.ensure(px * 48), "allocating the folded sums");
    rm::Params P = {};
    P.width = c->width;
    P.height = c->height;
    rm::launch_fold(P, c->accum.as<const float4>(), c->galb.as<const float4>(), c->gnorm.as<const float4>(),
                    c->comp.as<const float4>(), c->folded.as<float4>(), stream);
    RM_HIP(hipGetLastError(), "launching the fold of compensated sums");
    for (int k = 0; k < 3; ++k) out[k] = c->folded.as<const float4>() + (size_t)k * px;
    return RM_OK;
}

namespace {
// Every growable buffer of a context but its environment and work counter.
std::vector<DevBuf*> working_buffers(rm_context* c) {
    return {&c->accum,     &c->galb,        &c->gnorm,       &c->light_a,     &c->light_b,   &c->denoised,
            &c->guides,    &c->albedo_frame, &c->splats[0],  &c->splats[1],   &c->gathered[0], &c->gathered[1],
            &c->photon_counts, &c->photon_starts, &c->photons, &c->gather_queue, &c->emitters, &c->spill, &c->side_spill, &c->rays,
            &c->hits,      &c->counters,    &c->staging,     &c->display,     &c->comp,      &c->folded,
            &c->wavefront, &c->split};
}
}  // namespace

uint64_t rmi::context_bytes(const rm_context* c) {
    uint64_t b = c->env.bytes + c->env_table.bytes + c->work.bytes;
    for (DevBuf* d : working_buffers(const_cast<rm_context*>(c))) b += d->bytes;
    return b;
}

int rmi::trim(rm_context* c) {
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    RM_HIP(hipStreamSynchronize(c->stream), "waiting for the stream");
    RM_HIP(hipStreamSynchronize(c->side), "waiting for the side stream");
    RM_HIP(hipStreamSynchronize(c->post), "waiting for the denoiser's stream");
    for (DevBuf* d : working_buffers(c)) d->release();
    c->emitters_host.clear();
    c->capacity = 0;
    c->width = c->height = 0;
    c->comp_live = false;
    c->denoise_pending = false;
    c->rendered = false;
    return RM_OK;
}

// The tree traversals walk: the FP32 nodes, or when quantized nodes are
// asked for, the 4-wide quantized tree, else the binary quantized one, else
// (a tree that can't be quantized) the FP32 nodes.
int rmi::tree_of(const rm_context* c, const rm_scene* s) {
    if (!(c->options[RM_OPTION_PRECISION] & RM_PRECISION_QUANTIZED_BVH)) return rm::TREE_FP32;
    if (s->wnodes.ptr) return rm::TREE_WIDE;
    return s->qnodes.ptr ? rm::TREE_QUANTIZED : rm::TREE_FP32;
}

namespace {
using rmi::tree_of;

int upload(rm_context* c, DevBuf& b, const void* src, size_t bytes, const char* what) {
    size_t alloc = bytes < 16 ? 16 : bytes;
    RM_HIP(b.ensure(alloc), what);
    if (bytes) RM_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c->stream), what);
    return RM_OK;
}

int ensure_spill(rm_context* c, uint32_t slots) {
    RM_HIP(c->spill.ensure((size_t)slots * rm::SPILL_ENTRIES * sizeof(uint2)), "allocating the traversal spill");
    return RM_OK;
}

}  // namespace

extern "C" {

const char* rm_last_error(void) { return g_error.c_str(); }

int rm_open(int device, rm_context** out) {
    if (!out) return fail(RM_ERR_INVALID, "rm_open needs somewhere to put the context.");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(RM_ERR_DEVICE, "No HIP device is available.");
    if (device < 0 || device >= count)
        return fail(RM_ERR_INVALID, "Device " + std::to_string(device) + " doesn't exist (" + std::to_string(count) +
                                        " available).");
    RM_HIP(hipSetDevice(device), "selecting the device");
    rm_context* c = new rm_context();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->resolved_ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->resolved_ev[1], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->denoised_pair_ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->denoised_pair_ev[1], hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->post, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->prepared_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->denoised_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->batch_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->side_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&c->tune_ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&c->tune_ev[1]);
    c->owner.device = device;
    c->owner.streams[0] = c->stream;
    c->owner.streams[1] = c->side;
    c->owner.streams[2] = c->post;
    for (DevBuf* b : working_buffers(c)) b->owner = &c->owner;
    for (DevBuf* b : {&c->env, &c->env_table, &c->work}) b->owner = &c->owner;
    if (e != hipSuccess) {
        for (hipEvent_t ev : c->tune_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (c->side_ev) (void)hipEventDestroy(c->side_ev);
        if (c->batch_ev) (void)hipEventDestroy(c->batch_ev);
        if (c->denoised_ev) (void)hipEventDestroy(c->denoised_ev);
        if (c->prepared_ev) (void)hipEventDestroy(c->prepared_ev);
        if (c->post) (void)hipStreamDestroy(c->post);
        for (hipEvent_t ev : c->resolved_ev)
            if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : c->denoised_pair_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (c->join) (void)hipEventDestroy(c->join);
        if (c->side) (void)hipStreamDestroy(c->side);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "creating the streams");
    }
    // A 1×1 black environment and an empty table stand in until one is given.
    uint16_t black[4] = {0, 0, 0, 0x3c00};
    float table[1] = {0.0f};
    int r = rm_environment(c, black, 1, 1, table, 1, 0);
    if (r != RM_OK) {
        rm_close(c);
        return r;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->compute_units = prop.multiProcessorCount;
    e = static_scratch_of(device, &c->static_scratch);
    if (e != hipSuccess) {
        rm_close(c);
        return hip_fail(e, "finding the static raycast scratch");
    }
    if (c->work.ensure(256) != hipSuccess) {
        rm_close(c);
        return fail(RM_ERR_MEMORY, "allocating the work counter");
    }
    *out = c;
    return RM_OK;
}

int rm_set_option(rm_context* c, int option, int64_t value) {
    if (!c) return fail(RM_ERR_INVALID, "rm_set_option needs a context.");
    if (option < 0 || option >= RM_OPTION_COUNT) return fail(RM_ERR_INVALID, "Unknown option " + std::to_string(option) + ".");
    if (option == RM_OPTION_SCHEDULE && (value < RM_SCHEDULE_PIXEL || value > RM_SCHEDULE_SPLIT))
        return fail(RM_ERR_INVALID, "RM_OPTION_SCHEDULE is one of RM_SCHEDULE_PIXEL, _COMPACT, _PERSISTENT, "
                                    "_WAVEFRONT, _SORTED or _SPLIT.");
    if (option == RM_OPTION_SPLIT_DEPTH && (value < 0 || value > 8))
        return fail(RM_ERR_INVALID, "RM_OPTION_SPLIT_DEPTH is 1 to 8 bounces (0: the default, 2).");
    if (option == RM_OPTION_TUNE_CANDIDATES && (value < 0 || value >= (int64_t)1 << RM_TUNE_CANDIDATES))
        return fail(RM_ERR_INVALID, "RM_OPTION_TUNE_CANDIDATES is a mask of the " +
                                        std::to_string(RM_TUNE_CANDIDATES) + " candidates (0: all).");
    if (option == RM_OPTION_PERSISTENT_BLOCKS && (value < 0 || value > (1 << 20)))
        return fail(RM_ERR_INVALID, "RM_OPTION_PERSISTENT_BLOCKS is 0 (automatic) to 2^20.");
    if (option == RM_OPTION_WAVES && value != 0 && value != 4 && value != 5)
        return fail(RM_ERR_INVALID, "RM_OPTION_WAVES is 4 or 5 (0: the default, 4).");
    if (option == RM_OPTION_STACK_DEPTH && value != 0 && value != 8 && value != 16)
        return fail(RM_ERR_INVALID, "RM_OPTION_STACK_DEPTH is 8 or 16 (0: the default, 8).");
    if (option == RM_OPTION_BLOCK_SHAPE && value != RM_BLOCK_SQUARES && value != RM_BLOCK_ROWS)
        return fail(RM_ERR_INVALID, "RM_OPTION_BLOCK_SHAPE is RM_BLOCK_SQUARES or RM_BLOCK_ROWS.");
    if (option == RM_OPTION_BATCH && (value < 0 || value > 1))
        return fail(RM_ERR_INVALID, "RM_OPTION_BATCH is 0 or 1.");
    if (option == RM_OPTION_LIGHT_PATHS && (value < 0 || value > (1 << 28)))
        return fail(RM_ERR_INVALID, "RM_OPTION_LIGHT_PATHS is 0 (the reference's count) to 2^28 a frame.");
    if (option == RM_OPTION_AUTOTUNE && (value < 0 || value > 64))
        return fail(RM_ERR_INVALID, "RM_OPTION_AUTOTUNE is 0 (off) to 64 frames per candidate.");
    if (option == RM_OPTION_PRECISION &&
        (value & ~(int64_t)(RM_PRECISION_QUANTIZED_BVH | RM_PRECISION_FP16_GUIDES | RM_PRECISION_COMPENSATED)))
        return fail(RM_ERR_INVALID, "RM_OPTION_PRECISION takes RM_PRECISION_* flags.");
    c->options[option] = value;
    return RM_OK;
}

int rm_close(rm_context* c) {
    if (!c) return fail(RM_ERR_INVALID, "rm_close needs a context.");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->post) (void)hipStreamSynchronize(c->post);
    for (rm_scene* s : c->scenes) {
        s->release();
        delete s;
    }
    for (DevBuf* b : {&c->accum, &c->galb, &c->gnorm, &c->light_a, &c->light_b, &c->denoised, &c->splats[0], &c->splats[1],
                      &c->guides, &c->albedo_frame, &c->photon_counts, &c->photon_starts, &c->photons, &c->gather_queue, &c->emitters, &c->spill, &c->side_spill,
                      &c->gathered[0], &c->gathered[1], &c->env, &c->env_table, &c->rays, &c->hits, &c->counters,
                      &c->work, &c->staging, &c->display, &c->comp, &c->folded, &c->wavefront, &c->split})
        b->release();
    // The streams are idle: the blocks just given back need no events, which
    // must not outlive the streams they were recorded on.
    rmi::cache_forget(c->owner);
    for (int s = 0; s < RM_STAGE_COUNT; ++s)
        for (auto& p : c->timer.pending[s]) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    for (hipEvent_t e : c->timer.pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->tune_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->side_ev) (void)hipEventDestroy(c->side_ev);
    if (c->batch_ev) (void)hipEventDestroy(c->batch_ev);
    if (c->denoised_ev) (void)hipEventDestroy(c->denoised_ev);
    if (c->prepared_ev) (void)hipEventDestroy(c->prepared_ev);
    if (c->post) (void)hipStreamDestroy(c->post);
    for (hipEvent_t ev : c->resolved_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : c->denoised_pair_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RM_OK;
}

// The delta cut of a packed tree: at most DELTA_CUT_BOXES boxes whose union
// covers the box of every leaf holding a delta primitive (glass, or metal
// with roughness under 0.01, as rm::is_delta decides), each box the FP32
// min/max union of such leaf boxes, so exact. The slab test's entry distance
// only falls and its exit only grows as a box grows, so a ray that misses
// every cut box misses every such leaf's box, and the FP32 walk, which
// slab-tests a leaf's box before its primitives, never tests a delta
// primitive for it: its closest hit is not a delta surface. The photon
// gather's camera paths end at a first hit that isn't one, so it skips
// tracing them. The cut starts at the root's box and splits the box of
// largest area into its children's, while it has room. False (no cut) when
// the nodes do not form a tree or a delta leaf's box isn't finite.
constexpr size_t DELTA_CUT_BOXES = 16;
static bool delta_cut(const uint32_t* nw, uint32_t node_count, const uint32_t* pw, const uint32_t* mw,
                      std::vector<float>& out) {
    struct Box { float lo[3], hi[3]; bool any = false; };
    std::vector<Box> box(node_count);
    std::vector<uint8_t> seen(node_count, 0);
    auto node_box = [&](uint32_t i, Box& b) {
        for (int a = 0; a < 3; ++a) {
            std::memcpy(&b.lo[a], &nw[i * 8 + a], 4);
            std::memcpy(&b.hi[a], &nw[i * 8 + 4 + a], 4);
        }
    };
    auto is_delta = [&](uint32_t prim) {
        const uint32_t m = pw[prim * 16 + 1];
        const uint32_t kind = mw[m * 24 + 3];
        float roughness;
        std::memcpy(&roughness, &mw[m * 24 + 7], 4);
        return kind == rm::GLASS || (kind == rm::METAL && roughness < 0.01f);
    };
    // Children before parents: a node is finished when popped the second time.
    std::vector<std::pair<uint32_t, bool>> stack{{0u, false}};
    while (!stack.empty()) {
        const auto [i, children_done] = stack.back();
        stack.pop_back();
        const uint32_t first = nw[i * 8 + 3], count = nw[i * 8 + 7];
        Box& b = box[i];
        if (!children_done) {
            if (seen[i]++) return false;
            if (count == 0) {
                stack.push_back({i, true});
                stack.push_back({first, false});
                stack.push_back({first + 1, false});
                continue;
            }
            for (uint32_t k = first; k < first + count && !b.any; ++k) b.any = pw[k * 16] <= rm::TRIANGLE && is_delta(k);
            if (b.any) {
                node_box(i, b);
                for (int a = 0; a < 3; ++a)
                    if (!std::isfinite(b.lo[a]) || !std::isfinite(b.hi[a])) return false;
            }
            continue;
        }
        for (uint32_t child : {first, first + 1}) {
            const Box& c = box[child];
            if (!c.any) continue;
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = b.any ? std::fmin(b.lo[a], c.lo[a]) : c.lo[a];
                b.hi[a] = b.any ? std::fmax(b.hi[a], c.hi[a]) : c.hi[a];
            }
            b.any = true;
        }
    }
    std::vector<uint32_t> cut;
    if (box[0].any) cut.push_back(0);
    auto area = [&](uint32_t i) {
        const Box& b = box[i];
        const double x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
        return x * y + y * z + z * x;
    };
    for (;;) {
        size_t split = cut.size();
        for (size_t k = 0; k < cut.size(); ++k)
            if (nw[cut[k] * 8 + 7] == 0 && (split == cut.size() || area(cut[k]) > area(cut[split]))) split = k;
        if (split == cut.size()) break;
        const uint32_t first = nw[cut[split] * 8 + 3];
        const size_t grows = (box[first].any ? 1 : 0) + (box[first + 1].any ? 1 : 0) - 1;
        if (cut.size() + grows > DELTA_CUT_BOXES) break;
        cut.erase(cut.begin() + split);
        for (uint32_t child : {first, first + 1})
            if (box[child].any) cut.push_back(child);
    }
    out.clear();
    for (uint32_t i : cut) {
        const Box& b = box[i];
        out.insert(out.end(), {b.lo[0], b.lo[1], b.lo[2], 0.0f, b.hi[0], b.hi[1], b.hi[2], 0.0f});
    }
    return true;
}

int rm_scene_upload(rm_context* c, const void* primitives, uint32_t primitive_slots, uint32_t primitive_count,
                    const void* nodes, uint32_t node_count, const void* materials, uint32_t material_count,
                    const float* attributes, uint32_t attribute_vec4s, const uint32_t* lights, uint32_t lights_len,
                    rm_scene** out) {
    if (!c || !out) return fail(RM_ERR_INVALID, "rm_scene_upload needs a context and an output.");
    if (!primitives || primitive_slots == 0 || primitive_count > primitive_slots)
        return fail(RM_ERR_INVALID, "A scene needs at least one primitive slot, and no more in the BVH than slots.");
    if (!nodes || node_count == 0) return fail(RM_ERR_INVALID, "A scene needs at least one BVH node.");
    if (!materials && material_count) return fail(RM_ERR_INVALID, "Materials are missing.");
    // Every node's children and leaf ranges must stay inside the buffers the
    // kernels index.
    const uint32_t* nw = static_cast<const uint32_t*>(nodes);
    for (uint32_t i = 0; i < node_count; ++i) {
        uint32_t first = nw[i * 8 + 3], count = nw[i * 8 + 7];
        bool ok = count > 0 ? (uint64_t)first + count <= primitive_slots : (uint64_t)first + 1 < node_count;
        if (!ok) return fail(RM_ERR_INVALID, "BVH node " + std::to_string(i) + " points outside the scene.");
    }
    const uint32_t* pw = static_cast<const uint32_t*>(primitives);
    for (uint32_t i = 0; i < primitive_slots; ++i) {
        uint32_t kind = pw[i * 16], mat = pw[i * 16 + 1], attr = pw[i * 16 + 3];
        if (kind <= 3 && mat >= material_count)
            return fail(RM_ERR_INVALID, "Primitive " + std::to_string(i) + " has no material.");
        if (kind == 3 && attr != 0xffffffffu && (uint64_t)attr + 4 > attribute_vec4s)
            return fail(RM_ERR_INVALID, "Triangle " + std::to_string(i) + " has no attribute block.");
    }
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    rm_scene* s = new rm_scene();
    s->primitive_slots = primitive_slots;
    s->primitive_count = primitive_count;
    s->node_count = node_count;
    s->material_count = material_count;
    const float* mf = static_cast<const float*>(materials);
    for (uint32_t i = 0; i < material_count; ++i)
        if (mf[i * 24 + 19] > 0.0f) s->dispersive = true;
    int r = RM_OK;
    if (r == RM_OK) r = upload(c, s->prims, primitives, (size_t)primitive_slots * 64, "uploading primitives");
    if (r == RM_OK) r = upload(c, s->nodes, nodes, (size_t)node_count * 32, "uploading BVH nodes");
    if (r == RM_OK) {
        // The quantized trees (binary and 4-wide), for RM_PRECISION_QUANTIZED_BVH;
        // a tree they can't enclose (boxes that aren't finite) keeps only the
        // FP32 form, without an error.
        std::vector<uint32_t> q((size_t)node_count * 8);
        const std::string before = rm_last_error();
        if (rm_bvh_quantize(nodes, node_count, q.data()) == RM_OK)
            r = upload(c, s->qnodes, q.data(), q.size() * 4, "uploading the quantized BVH");
        std::vector<uint32_t> w((size_t)node_count * 16);
        uint32_t wide = 0;
        if (r == RM_OK && rm_bvh_quantize_wide(nodes, node_count, w.data(), node_count, &wide) == RM_OK && wide > 0)
            r = upload(c, s->wnodes, w.data(), (size_t)wide * 64, "uploading the wide quantized BVH");
        if (r == RM_OK) fail(RM_OK, before);
        if (r == RM_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(RM_ERR_DEVICE, "uploading a scene");
    }
    if (r == RM_OK) {
        std::vector<float> cut;
        if (delta_cut(nw, node_count, pw, static_cast<const uint32_t*>(materials), cut)) {
            const uint32_t boxes = (uint32_t)(cut.size() / 8);
            if (cut.empty()) cut.assign(8, 0.0f);   // an address for an empty cut
            r = upload(c, s->delta, cut.data(), cut.size() * 4, "uploading the delta cut");
            if (r == RM_OK) s->delta_count = boxes;
        }
    }
    if (r == RM_OK) r = upload(c, s->mats, materials, (size_t)material_count * 96, "uploading materials");
    if (r == RM_OK) r = upload(c, s->attr, attributes, (size_t)attribute_vec4s * 16, "uploading attributes");
    if (r == RM_OK) r = upload(c, s->lights, lights, (size_t)lights_len * 4, "uploading the lights table");
    if (r == RM_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(RM_ERR_DEVICE, "uploading a scene");
    if (r == RM_OK && lights) s->lights_host.assign(lights, lights + lights_len);
    if (r == RM_OK) {
        s->kinds_host.resize(primitive_slots);
        for (uint32_t i = 0; i < primitive_slots; ++i) s->kinds_host[i] = (uint8_t)(pw[i * 16] < 255u ? pw[i * 16] : 255u);
    }
    if (r != RM_OK) {
        s->release();
        delete s;
        return r;
    }
    c->scenes.push_back(s);
    *out = s;
    return RM_OK;
}

int rm_scene_release(rm_context* c, rm_scene* s) {
    if (!c || !s) return fail(RM_ERR_INVALID, "rm_scene_release needs a context and a scene.");
    for (size_t i = 0; i < c->scenes.size(); ++i)
        if (c->scenes[i] == s) {
            (void)hipSetDevice(c->device);
            (void)hipStreamSynchronize(c->stream);
            s->release();
            delete s;
            c->scenes.erase(c->scenes.begin() + i);
            return RM_OK;
        }
    return fail(RM_ERR_INVALID, "That scene doesn't belong to this context.");
}

int rm_scene_textures(rm_context* c, rm_scene* s, int normal_maps, const uint8_t* rgba, uint32_t width,
                      uint32_t height, uint32_t layers) {
    if (!c || !s || (!rgba && layers)) return fail(RM_ERR_INVALID, "rm_scene_textures needs a context, scene and data.");
    if (layers && (width == 0 || height == 0)) return fail(RM_ERR_INVALID, "Texture layers need a size.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    DevBuf& b = normal_maps ? s->nmaps : s->textures;
    int r = upload(c, b, rgba, (size_t)width * height * layers * 4, "uploading textures");
    if (r != RM_OK) return r;
    RM_HIP(hipStreamSynchronize(c->stream), "uploading textures");
    if (normal_maps) { s->nmap_w = width; s->nmap_h = height; s->nmap_layers = layers; }
    else { s->tex_w = width; s->tex_h = height; s->tex_layers = layers; }
    return RM_OK;
}

int rm_environment(rm_context* c, const uint16_t* rgba_half, uint32_t width, uint32_t height, const float* table,
                   uint32_t table_width, uint32_t table_height) {
    if (!c || !rgba_half || !table || width == 0 || height == 0 || table_width == 0)
        return fail(RM_ERR_INVALID, "An environment needs texels and a table.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    int r = upload(c, c->env, rgba_half, (size_t)width * height * 8, "uploading the environment");
    if (r == RM_OK)
        r = upload(c, c->env_table, table, (size_t)table_width * (table_height + 1) * 4, "uploading the environment table");
    if (r != RM_OK) return r;
    RM_HIP(hipStreamSynchronize(c->stream), "uploading the environment");
    c->env_w = width;
    c->env_h = height;
    c->table_w = table_width;
    c->table_h = table_height;
    return RM_OK;
}

int rm_resize(rm_context* c, uint32_t width, uint32_t height) {
    if (!c || width == 0 || height == 0) return fail(RM_ERR_INVALID, "An image needs a width and height.");
    if ((uint64_t)width * height > (1ull << 28)) return fail(RM_ERR_INVALID, "That image is too large.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    uint32_t px = width * height;
    if (px > c->capacity) {
        // Earlier frames on any of the three streams may still use the buffers.
        RM_HIP(hipStreamSynchronize(c->stream), "resizing");
        RM_HIP(hipStreamSynchronize(c->side), "resizing");
        RM_HIP(hipStreamSynchronize(c->post), "resizing");
        size_t bytes = (size_t)px * 16;
        for (DevBuf* b : {&c->accum, &c->galb, &c->gnorm, &c->light_a, &c->light_b, &c->denoised, &c->splats[0],
                          &c->splats[1], &c->guides}) {
            b->release();
            RM_HIP(b->ensure(bytes), "allocating per-pixel buffers");
        }
        RM_HIP(hipMemsetAsync(c->accum.ptr, 0, bytes, c->stream), "clearing the accumulation");
        RM_HIP(hipMemsetAsync(c->galb.ptr, 0, bytes, c->stream), "clearing the guides");
        RM_HIP(hipMemsetAsync(c->gnorm.ptr, 0, bytes, c->stream), "clearing the guides");
        RM_HIP(hipMemsetAsync(c->splats[0].ptr, 0, bytes, c->stream), "clearing the splats");
        RM_HIP(hipMemsetAsync(c->splats[1].ptr, 0, bytes, c->stream), "clearing the splats");
        // Light tracing, on the side stream, adds to the cleared splats.
        RM_HIP(hipStreamSynchronize(c->stream), "clearing the splats");
        c->capacity = px;
        c->comp_live = false;
    }
    int r = ensure_spill(c, px);
    if (r != RM_OK) return r;
    c->width = width;
    c->height = height;
    return RM_OK;
}

// A frame in three parts around its path pass, so that rm_frame_batch can
// give several contexts' path passes one launch.
struct Frame {
    const rm_scene* s;
    rm::Params P;
    rm::DevScene S;
    rm::Features f;
    bool caustics;
    float4* comp;
    uint32_t px;
    uint32_t pair = 0;   // the splat and gather pair of a frame with light tracing
    // Set by frame_alloc: light tracing's photons and spill lanes; the path
    // pass's spill lanes, persistent grid, staging records, wavefront state.
    uint32_t photons = 0, capacity = 0, side_slots = 0;
    uint32_t slots = 0, persistent = 0;
    float4* staging = nullptr;
    bool wavefront = false, sorted = false;
    uint32_t* split = nullptr;
    uint32_t split_blocks = 0;
};

// The photon gather's queue for px pixels: a 96-byte slot each, then its length.
static size_t gather_queue_bytes(size_t px) { return px * rm::GATHER_SLOT_VEC4S * 16 + 256; }

// The checks of rm_frame's arguments, before anything is enqueued.
static int frame_check(rm_context* c, const rm_scene* s, const void* params, const float* emitters,
                       uint32_t emitter_vec4s, Frame& F) {
    if (!c || !s || !params) return fail(RM_ERR_INVALID, "rm_frame needs a context, a scene and parameters.");
    rm::Params& P = F.P;
    std::memcpy(&P, params, sizeof P);
    if (c->width == 0) return fail(RM_ERR_STATE, "rm_frame before rm_resize.");
    if (P.width != c->width || P.height != c->height)
        return fail(RM_ERR_INVALID, "The parameters' size differs from rm_resize's.");
    if (P.samples_per_frame == 0) return fail(RM_ERR_INVALID, "A frame needs at least one sample per pixel.");
    if ((uint64_t)P.light_count * 2 + P.volume_count > s->lights_host.size())
        return fail(RM_ERR_INVALID, "The lights table is shorter than the parameters' light and volume counts.");
    for (uint32_t i = 0; i < P.light_count + P.volume_count; ++i)
        if (s->lights_host[i] >= s->primitive_slots)
            return fail(RM_ERR_INVALID, "The lights table names a primitive the scene doesn't have.");
    const bool caustics = P.caustics != 0u;
    if (caustics && (!emitters || emitter_vec4s < 2))
        return fail(RM_ERR_INVALID, "Light tracing needs the emitters.");
    if (caustics) {
        const uint32_t* ew = reinterpret_cast<const uint32_t*>(emitters);
        uint64_t need = 1ull + ew[0] + ew[1];
        // Light tracing aims every photon at one of the targets and samples
        // a source as a quad or a sphere (or the sun, NONE).
        if (ew[0] == 0 || ew[1] == 0)
            return fail(RM_ERR_INVALID, "Light tracing needs at least one emitter and one target.");
        if (need > emitter_vec4s)
            return fail(RM_ERR_INVALID, "The emitters buffer is shorter than its header says.");
        for (uint32_t k = 0; k < ew[0]; ++k) {
            uint32_t prim = ew[(1 + k) * 4 + 1];
            if (prim != 0xffffffffu && prim >= s->primitive_slots)
                return fail(RM_ERR_INVALID, "An emitter names a primitive the scene doesn't have.");
            if (prim != 0xffffffffu && s->kinds_host[prim] != rm::QUAD && s->kinds_host[prim] != rm::SPHERE)
                return fail(RM_ERR_INVALID, "An emitter is neither a quad nor a sphere.");
        }
    }
    F.s = s;
    F.caustics = caustics;
    return RM_OK;
}

// Every allocation a frame needs, before anything of it is enqueued: a
// frame that fails here leaves the context as it was (no light tracing
// queued into a splat pair that no resolve would then clear).
// path_buffers: also the buffers of the context's own path-pass schedule
// (a batched frame's path pass runs in the batch's kernel instead).
static int frame_alloc(rm_context* c, const void* emitters, uint32_t emitter_vec4s, Frame& F, bool path_buffers) {
    const rm_scene* s = F.s;
    const rm::Params& P = F.P;
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    F.S = dev_scene(c, s);
    F.f = {P.volume_count > 0u, s->tex_layers + s->nmap_layers > 0, P.env_strength > 0.0f && c->table_h > 0,
           s->dispersive, tree_of(c, s)};
    const uint32_t px = F.px = c->width * c->height;
    if (F.caustics) {
        // Light paths and photon slots: the reference's count and fixed
        // capacity, or the caller's budget for both.
        const int64_t budget = c->options[RM_OPTION_LIGHT_PATHS];
        const uint32_t count = budget > 0 ? (uint32_t)budget : rm::photon_count(P);
        F.photons = count;
        F.capacity = budget > 0 ? count : rm::PHOTON_CAPACITY;
        // Traversal spill lanes: light tracing's, or the gather's (a lane a
        // pixel), which runs after it on the same stream.
        const uint32_t light_lanes = count < rm::LIGHT_LANES ? count : rm::LIGHT_LANES;
        F.side_slots = light_lanes > px ? light_lanes : px;
        const uint32_t pair = F.pair = c->pair;
        const size_t emitter_bytes = (size_t)emitter_vec4s * 16;
        const size_t spill_bytes = (size_t)F.side_slots * rm::SPILL_ENTRIES * sizeof(uint2);
        // Growing a buffer frees the old one, and new emitters overwrite the
        // old: earlier frames' light tracing may still be using either.
        const bool same_emitters = c->emitters_host.size() == emitter_bytes &&
                                   std::memcmp(c->emitters_host.data(), emitters, emitter_bytes) == 0;
        const size_t queue_bytes = gather_queue_bytes(px);
        if (!same_emitters || c->photon_counts.bytes == 0 || c->photon_starts.bytes == 0 ||
            c->photons.bytes < (size_t)F.capacity * 96 || c->gather_queue.bytes < queue_bytes ||
            c->gathered[pair].bytes < (size_t)px * 16 || c->side_spill.bytes < spill_bytes) {
            RM_HIP(hipStreamSynchronize(c->side), "waiting for light tracing");
        }
        RM_HIP(c->gather_queue.ensure(queue_bytes), "allocating the gather queue");
        RM_HIP(c->photon_counts.ensure((size_t)(rm::PHOTON_TABLE + 1) * 4), "allocating the photon grid");
        RM_HIP(c->photon_starts.ensure((size_t)(rm::PHOTON_TABLE + 2 * rm::PHOTON_GROUPS) * 4), "allocating the photon grid");
        RM_HIP(c->photons.ensure((size_t)F.capacity * 96), "allocating photons");
        RM_HIP(c->gathered[pair].ensure((size_t)px * 16), "allocating the gathered light");
        RM_HIP(c->side_spill.ensure(spill_bytes), "allocating the side traversal spill");
        if (!same_emitters) {
            c->emitters_host.clear();
            RM_HIP(c->emitters.ensure(emitter_bytes), "allocating emitters");
            RM_HIP(hipMemcpy(c->emitters.ptr, emitters, emitter_bytes, hipMemcpyHostToDevice), "uploading emitters");
            const uint8_t* e = reinterpret_cast<const uint8_t*>(emitters);
            c->emitters_host.assign(e, e + emitter_bytes);
        }
    }
    const bool compensated = (c->options[RM_OPTION_PRECISION] & RM_PRECISION_COMPENSATED) != 0;
    if (compensated) {
        RM_HIP(c->comp.ensure((size_t)px * 48), "allocating the compensation terms");
        RM_HIP(c->folded.ensure((size_t)px * 48), "allocating the folded sums");
    }
    if (P.denoise != 0u) RM_HIP(c->albedo_frame.ensure((size_t)px * 16), "allocating the denoiser's albedo");

    // The path pass's own buffers, by schedule.
    F.slots = px;
    F.persistent = 0;
    F.staging = nullptr;
    F.wavefront = false;
    F.sorted = false;
    F.split = nullptr;
    F.split_blocks = 0;
    const int64_t schedule = path_buffers ? c->options[RM_OPTION_SCHEDULE] : RM_SCHEDULE_PIXEL;
    if (schedule == RM_SCHEDULE_PERSISTENT && P.view == 0u) {
        int64_t blocks = c->options[RM_OPTION_PERSISTENT_BLOCKS];
        if (blocks == 0)
            blocks = (int64_t)(c->compute_units > 0 ? c->compute_units : 1) * rm::main_resident_blocks_per_cu(F.f);
        F.persistent = (uint32_t)blocks;
        const uint32_t threads = F.persistent * rm::BLOCK_THREADS;
        int r = ensure_spill(c, threads > px ? threads : px);
        if (r != RM_OK) return r;
        F.slots = threads;
    } else if ((schedule == RM_SCHEDULE_COMPACT || schedule == RM_SCHEDULE_SORTED) && P.view == 0u) {
        F.sorted = schedule == RM_SCHEDULE_SORTED;
        const uint32_t threads = ((c->width + 15u) / 16u) * ((c->height + 15u) / 16u) * rm::BLOCK_THREADS;
        int r = ensure_spill(c, threads);
        if (r != RM_OK) return r;
        RM_HIP(c->staging.ensure((size_t)threads * 48), "allocating the path staging records");
        F.staging = c->staging.as<float4>();
        F.slots = threads;
    } else if (schedule == RM_SCHEDULE_WAVEFRONT && P.view == 0u) {
        // Every stage's grid covers all paths, a thread per queue entry.
        const uint32_t threads = (px + rm::BLOCK_THREADS - 1) / rm::BLOCK_THREADS * rm::BLOCK_THREADS;
        int r = ensure_spill(c, threads);
        if (r != RM_OK) return r;
        RM_HIP(c->wavefront.ensure(rm::wavefront_bytes(px)), "allocating the wavefront path state");
        F.wavefront = true;
        F.slots = threads;
    } else if (schedule == RM_SCHEDULE_SPLIT && P.view == 0u && P.samples_per_frame == 1u) {
        // The tail's grid, with no more blocks than the queue could fill. A
        // lane's stack spill is its pixel's in the head and its grid slot in
        // the tail.
        F.split_blocks = std::min(rm::split_tail_blocks(c->compute_units, c->options[RM_OPTION_WAVES] == 5 ? 5 : 4),
                                  (px + rm::BLOCK_THREADS - 1) / rm::BLOCK_THREADS);
        const uint32_t threads = F.split_blocks * rm::BLOCK_THREADS;
        const uint32_t lanes = threads > px ? threads : px;
        int r = ensure_spill(c, lanes);
        if (r != RM_OK) return r;
        RM_HIP(c->split.ensure(rm::split_bytes(px)), "allocating the split schedule's path queue");
        F.split = c->split.as<uint32_t>();
        F.slots = lanes;
    }
    return RM_OK;
}

// The context's photon map for a frame of `capacity` photon slots (frame_alloc
// sized its buffers): the deposits, then their sorted positions and powers.
static rm::PhotonMap photon_map(rm_context* c, uint32_t capacity) {
    rm::PhotonMap M;
    M.counts = c->photon_counts.as<uint32_t>();
    M.starts = c->photon_starts.as<uint32_t>();
    M.deposits = c->photons.as<float4>();
    M.positions = M.deposits + (size_t)capacity * 3;
    M.powers = M.positions + capacity;
    M.capacity = capacity;
    return M;
}

// A frame's light tracing as one job of a batched launch (rm_frame_batch).
static rm::LightJob light_job(rm_context* c, const Frame& F) {
    return rm::LightJob{F.S, F.P, c->emitters.as<const float4>(), c->splats[F.pair].as<uint32_t>(),
                        photon_map(c, F.capacity), c->side_spill.as<uint32_t>(), F.side_slots, F.photons};
}

// After light tracing, on the side stream: the photon map's sort and the
// gather, then the join the path stream's resolve waits for.
static int light_end(rm_context* c, const Frame& F) {
    const rm::PhotonMap M = photon_map(c, F.capacity);
    const rm::GatherQueue Q{c->gather_queue.as<float4>(),
                            reinterpret_cast<uint32_t*>(c->gather_queue.as<char>() + gather_queue_bytes(F.px) - 256),
                            F.px};
    {
        Timed t(c, RM_STAGE_GATHER, c->side);
        rm::launch_photon_sort(M, c->side);
        rm::launch_photon_gather(F.S, F.P, M, Q, c->gathered[F.pair].as<float4>(), c->side_spill.as<uint32_t>(),
                                 F.side_slots, F.f, rm::side_waves(c->options[RM_OPTION_WAVES] == 5 ? 5 : 4), c->side);
    }
    RM_HIP(hipEventRecord(c->join, c->side), "joining the side stream");
    RM_HIP(hipGetLastError(), "launching light tracing");
    return RM_OK;
}

// Everything before the path pass: light tracing and the photon gather on
// the side stream, the compensation terms. batched_light: stop after the
// cell counts' clear; rm_frame_batch traces the batch's light paths in one
// launch, then runs light_end for each frame.
static int frame_begin(rm_context* c, Frame& F, bool batched_light = false) {
    const rm::Params& P = F.P;
    const rm::DevScene& S = F.S;
    const rm::Features& f = F.f;
    // Light tracing and the photon gather depend on the scene and the
    // parameters only, so they run on the side stream beside the path pass;
    // k_resolve (or k_finish) then adds splats and gathered light, after the
    // path pass, in the reference's order, once the join says they are done.
    // The side stream reads nothing the path stream writes but the splats and
    // gathered light, double-buffered by frame: it waits only for the resolve
    // that last read this frame's pair, so it can start while the previous
    // frame's path pass still runs. It also waits for the denoiser of the
    // frame that last used the pair, which ends some way into the previous
    // frame's path pass: started at that pass's beginning instead, light
    // tracing and the gather met it at full occupancy in about one run in
    // three (1.32 against 1.48 G samples/s on the 1080p sphere frame); with
    // the wait every run settles near 1.42. Scenes, environments and textures
    // are uploaded synchronously, and the emitters too, only when they change
    // and once the side stream has drained.
    if (F.caustics) {
        const uint32_t pair = F.pair;
        c->pair = pair ^ 1u;
        const rm::PhotonMap M = photon_map(c, F.capacity);
        RM_HIP(hipStreamWaitEvent(c->side, c->resolved_ev[pair], 0), "ordering light tracing");
        if (c->denoised_pair_live[pair])
            RM_HIP(hipStreamWaitEvent(c->side, c->denoised_pair_ev[pair], 0), "ordering light tracing");
        {
            Timed t(c, RM_STAGE_CLEAR, c->side);
            rm::launch_fill_u32(M.counts, 0u, rm::PHOTON_TABLE + 1, c->side);
        }
        if (!batched_light) {
            {
                Timed t(c, RM_STAGE_LIGHT_TRACE, c->side);
                rm::launch_light_trace(S, P, c->emitters.as<const float4>(), c->splats[pair].as<uint32_t>(), M,
                                       c->side_spill.as<uint32_t>(), F.side_slots, F.photons, f,
                                       rm::side_waves(c->options[RM_OPTION_WAVES] == 5 ? 5 : 4), c->side);
            }
            int r = light_end(c, F);
            if (r != RM_OK) return r;
        }
    }

    // Compensated sums: the error terms start at zero with the image (the
    // path pass writes them on a first frame), or when compensation starts
    // part-way through an accumulation.
    const bool compensated = (c->options[RM_OPTION_PRECISION] & RM_PRECISION_COMPENSATED) != 0;
    F.comp = nullptr;
    if (compensated) {
        F.comp = c->comp.as<float4>();
        if (!c->comp_live && P.sample_offset != 0u)
            rm::launch_fill_u32(c->comp.as<uint32_t>(), 0u, (size_t)F.px * 12, c->stream);
    }
    c->comp_live = compensated;
    return RM_OK;
}

// The wavefront path state of a frame of px pixels, carved out of the
// context's wavefront buffer (frame_alloc sized it).
static rm::Wavefront carve_wavefront(rm_context* c, uint32_t px) {
    rm::Wavefront W;
    char* base = c->wavefront.as<char>();
    auto carve = [&](size_t bytes) {
        char* p = base;
        base += (bytes + 255) & ~(size_t)255;
        return p;
    };
    W.n = px;
    W.state[0] = reinterpret_cast<uint32_t*>(carve((size_t)px * rm::WF_STATE_WORDS * 4));
    W.state[1] = reinterpret_cast<uint32_t*>(carve((size_t)px * rm::WF_STATE_WORDS * 4));
    W.hits = reinterpret_cast<uint4*>(carve((size_t)px * 16));
    W.nee = reinterpret_cast<uint32_t*>(carve((size_t)px * rm::WF_NEE_WORDS * 4));
    W.done = reinterpret_cast<uint32_t*>(carve((size_t)px * rm::WF_DONE_WORDS * 4));
    W.sums = reinterpret_cast<float4*>(carve((size_t)px * 48));
    W.counts = reinterpret_cast<uint32_t*>(carve(16));
    return W;
}

// The path pass under the context's schedule, into the buffers frame_alloc sized.
static int frame_path(rm_context* c, Frame& F) {
    const rm::Params& P = F.P;
    const rm::DevScene& S = F.S;
    const rm::Features& f = F.f;
    const uint32_t px = F.px;
    if (F.persistent) rm::launch_fill_u32(c->work.as<uint32_t>(), 0u, 1, c->stream);
    if (F.wavefront) {
        const rm::Wavefront W = carve_wavefront(c, px);
        Timed t(c, RM_STAGE_TRACE);
        rm::LaunchWavefront L{S, P, W, c->accum.as<float4>(), c->galb.as<float4>(), c->gnorm.as<float4>(), F.comp,
                              c->spill.as<uint32_t>(), F.slots, F.slots / rm::BLOCK_THREADS, f.tree};
        rm::launch_wavefront(L, f, c->stream);
    } else {
        Timed t(c, RM_STAGE_TRACE);
        rm::LaunchMain L{S, P, c->accum.as<float4>(), c->galb.as<float4>(), c->gnorm.as<float4>(), F.comp,
                         c->spill.as<uint32_t>(), F.slots, F.persistent, c->work.as<uint32_t>(), F.staging, f.tree,
                         c->options[RM_OPTION_WAVES] == 5 ? 5 : 4, F.sorted,
                         c->options[RM_OPTION_STACK_DEPTH] == 16 ? 16 : 8, (int)c->options[RM_OPTION_BLOCK_SHAPE],
                         F.split, c->options[RM_OPTION_SPLIT_DEPTH] > 0 ? (uint32_t)c->options[RM_OPTION_SPLIT_DEPTH] : 2u,
                         F.split_blocks};
        rm::launch_main(L, f, c->stream);
    }
    RM_HIP(hipGetLastError(), "launching the trace");
    return RM_OK;
}

// The denoiser's passes are done on the post stream (they were enqueued
// there, or the post stream waits for a batch that ran them): what later
// frames and reads wait for.
static int denoise_recorded(rm_context* c, const Frame& F) {
    RM_HIP(hipEventRecord(c->denoised_ev, c->post), "ordering the denoiser");
    if (F.caustics) {
        RM_HIP(hipEventRecord(c->denoised_pair_ev[F.pair], c->post), "ordering light tracing");
        c->denoised_pair_live[F.pair] = true;
    }
    c->denoise_pending = true;
    return RM_OK;
}

// A frame's denoiser passes as one job of a batched launch (rm_frame_batch).
static rm::DenoiseJob denoise_job(rm_context* c, const Frame& F) {
    return rm::DenoiseJob{F.P, c->guides.ptr, c->light_a.as<float4>(), c->light_b.as<float4>(),
                          c->albedo_frame.as<const float4>(), c->denoised.as<float4>()};
}

// Everything after the path pass: the splat and gather resolve, the
// denoiser. batched_denoise: stop once prepare is enqueued; rm_frame_batch
// runs the batch's à-trous passes and compose, then denoise_recorded.
static int frame_end(rm_context* c, Frame& F, bool batched_denoise = false) {
    const rm::Params& P = F.P;
    const bool caustics = F.caustics;
    float4* comp = F.comp;

    // From the fourth sample on, with plain sums, resolve, guides and prepare
    // run as one pass (k_finish) timed as the prepare stage.
    const bool fused = P.denoise != 0u && comp == nullptr && P.total_samples >= 4u;
    if (caustics) RM_HIP(hipStreamWaitEvent(c->stream, c->join, 0), "joining the side stream");
    if (caustics && !fused) {
        {
            Timed t(c, RM_STAGE_RESOLVE);
            rm::launch_resolve(P, c->splats[F.pair].as<uint4>(), c->gathered[F.pair].as<const float4>(),
                               c->accum.as<float4>(), comp, c->stream);
        }
        RM_HIP(hipEventRecord(c->resolved_ev[F.pair], c->stream), "ordering light tracing");
        RM_HIP(hipGetLastError(), "launching the splat resolve");
    }

    if (P.denoise != 0u) {
        const bool fp16 = (c->options[RM_OPTION_PRECISION] & RM_PRECISION_FP16_GUIDES) != 0;
        float4* a = c->light_a.as<float4>();
        float4* b = c->light_b.as<float4>();
        // The previous frame's passes still read light_a/b, guides and the albedo.
        if (c->denoise_pending) RM_HIP(hipStreamWaitEvent(c->stream, c->denoised_ev, 0), "ordering the denoiser");
        if (fused) {
            Timed t(c, RM_STAGE_PREPARE);
            rm::launch_finish(P, caustics ? c->splats[F.pair].as<uint4>() : nullptr,
                              caustics ? c->gathered[F.pair].as<const float4>() : nullptr, c->accum.as<float4>(),
                              c->galb.as<const float4>(), c->gnorm.as<const float4>(), c->guides.ptr, fp16, a,
                              c->albedo_frame.as<float4>(), c->stream);
            if (caustics) RM_HIP(hipEventRecord(c->resolved_ev[F.pair], c->stream), "ordering light tracing");
        } else {
            Timed t(c, RM_STAGE_PREPARE);
            const float4* sums[3];
            int r = rmi::current_sums(c, c->stream, sums);
            if (r != RM_OK) return r;
            rm::launch_guides(P, sums[2], c->guides.ptr, fp16, c->stream);
            rm::launch_prepare(P, sums[0], sums[1], c->guides.ptr, fp16, a, c->albedo_frame.as<float4>(), c->stream);
        }
        RM_HIP(hipEventRecord(c->prepared_ev, c->stream), "ordering the denoiser");
        if (!batched_denoise) {
            RM_HIP(hipStreamWaitEvent(c->post, c->prepared_ev, 0), "ordering the denoiser");
            const int steps[5] = {1, 2, 4, 8, 16};
            float4* in = a;
            float4* out = b;
            for (int k = 0; k < 5; ++k) {
                Timed t(c, RM_STAGE_ATROUS, c->post);
                rm::launch_atrous(P, steps[k], c->guides.ptr, fp16, in, out, c->post);
                std::swap(in, out);
            }
            {
                Timed t(c, RM_STAGE_COMPOSE, c->post);
                rm::launch_compose(P, c->albedo_frame.as<const float4>(), in, c->denoised.as<float4>(), c->post);
            }
            int r = denoise_recorded(c, F);
            if (r != RM_OK) return r;
        }
        RM_HIP(hipGetLastError(), "launching the denoiser");
    }
    c->last_params = P;
    c->rendered = true;
    return RM_OK;
}


int rm_frame(rm_context* c, const rm_scene* s, const void* params, const float* emitters, uint32_t emitter_vec4s) {
    Frame F;
    int r = frame_check(c, s, params, emitters, emitter_vec4s, F);
    if (r == RM_OK) r = frame_alloc(c, emitters, emitter_vec4s, F, true);
    if (r == RM_OK) r = frame_begin(c, F);
    if (r == RM_OK) r = frame_path(c, F);
    if (r == RM_OK) r = frame_end(c, F);
    return r;
}

namespace {
// The cross-job wavefront pass's launch sequences, captured in hipGraphs, one
// for each shape (device, features, tree, samples, bounce rounds, grid), with
// the device job table its launches read and an event that passes once the
// last replay has read it.
struct WfGraph {
    int device;
    uint64_t key;
    hipGraphExec_t exec;
    rm::WfTable* table;
    hipEvent_t done;
};
std::mutex g_graph_lock;
std::vector<WfGraph> g_graphs;
constexpr size_t MAX_GRAPHS = 32;

void drop_graph(const WfGraph& g) {
    (void)hipEventSynchronize(g.done);
    (void)hipGraphExecDestroy(g.exec);
    (void)hipFree(g.table);
    (void)hipEventDestroy(g.done);
}

// Captures the launch sequence of shape B on `stream` into a new graph.
int capture_graph(hipStream_t stream, int device, uint64_t key, const rm::WfBatchShape& B, const rm::Features& u,
                  WfGraph& out) {
    out = WfGraph{device, key, nullptr, nullptr, nullptr};
    RM_HIP(hipMalloc((void**)&out.table, sizeof(rm::WfTable)), "allocating a batch's job table");
    hipError_t e = hipEventCreateWithFlags(&out.done, hipEventDisableTiming);
    hipGraph_t graph = nullptr;
    if (e == hipSuccess) e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rm::launch_wavefront_batch(out.table, B, u, stream);
        hipError_t launched = hipGetLastError();
        e = hipStreamEndCapture(stream, &graph);
        if (e == hipSuccess) e = launched;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&out.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        if (out.exec) (void)hipGraphExecDestroy(out.exec);
        if (out.done) (void)hipEventDestroy(out.done);
        (void)hipFree(out.table);
        out = WfGraph{};
        return hip_fail(e, "capturing the batched wavefront pass");
    }
    return RM_OK;
}

// The batch's wavefront path passes: its job table uploaded on the first
// context's stream, then the graph of its shape replayed there.
int wavefront_batch(rm_context* run, rm_context* const* ctxs, const Frame* F, uint32_t n, const rm::Features& u) {
    rm::WfTable T;
    std::memset(&T, 0, sizeof T);
    rm::WfBatchShape B{0u, 0u, (uint32_t)(run->compute_units > 0 ? run->compute_units : 1) * 8u, u.tree};
    uint32_t chunks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        rm_context* c = ctxs[i];
        T.job[i] = rm::WfJob{F[i].S, F[i].P, carve_wavefront(c, F[i].px), c->accum.as<float4>(), c->galb.as<float4>(),
                             c->gnorm.as<float4>(), F[i].comp, c->spill.as<uint32_t>(), F[i].slots};
        T.begin[i] = chunks;
        chunks += (F[i].px + rm::BLOCK_THREADS - 1) / rm::BLOCK_THREADS;
        B.samples = std::max(B.samples, F[i].P.samples_per_frame);
        B.bounces = std::max(B.bounces, F[i].P.max_bounces + 1u);
    }
    T.begin[n] = chunks;
    T.n = n;
    const uint64_t features = (u.volumes ? 8u : 0u) | (u.textures ? 4u : 0u) | (u.environment ? 2u : 0u) |
                              (u.dispersion ? 1u : 0u);
    const uint64_t key = features | (uint64_t)(B.tree & 3) << 4 | (uint64_t)(B.samples & 0xffffu) << 8 |
                         (uint64_t)(B.bounces & 0xffffu) << 24 | (uint64_t)B.blocks << 40;
    std::lock_guard<std::mutex> guard(g_graph_lock);
    WfGraph* g = nullptr;
    for (WfGraph& x : g_graphs)
        if (x.device == run->device && x.key == key) g = &x;
    if (!g) {
        if (g_graphs.size() >= MAX_GRAPHS) {
            drop_graph(g_graphs.front());
            g_graphs.erase(g_graphs.begin());
        }
        WfGraph made;
        int r = capture_graph(run->stream, run->device, key, B, u, made);
        if (r != RM_OK) return r;
        g_graphs.push_back(made);
        g = &g_graphs.back();
    }
    // The last replay of this graph, on whichever stream, has read the table.
    RM_HIP(hipStreamWaitEvent(run->stream, g->done, 0), "ordering a batch's job table");
    RM_HIP(hipMemcpyAsync(g->table, &T, sizeof T, hipMemcpyHostToDevice, run->stream), "uploading a batch's job table");
    RM_HIP(hipGraphLaunch(g->exec, run->stream), "launching the batched wavefront pass");
    RM_HIP(hipEventRecord(g->done, run->stream), "ordering a batch's job table");
    return RM_OK;
}
}  // namespace

int rm_frame_batch(rm_context* const* ctxs, const rm_scene* const* scenes, const void* const* params,
                   const float* const* emitters, const uint32_t* emitter_vec4s, uint32_t n) {
    if (!ctxs || !scenes || !params || n == 0 || n > rm::MAX_BATCH)
        return fail(RM_ERR_INVALID, "rm_frame_batch needs 1 to " + std::to_string(rm::MAX_BATCH) +
                                        " contexts, scenes and parameter blocks.");
    Frame F[rm::MAX_BATCH];
    for (uint32_t i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(RM_ERR_INVALID, "rm_frame_batch needs a context for every frame.");
        for (uint32_t k = 0; k < i; ++k)
            if (ctxs[k] == ctxs[i]) return fail(RM_ERR_INVALID, "A batch takes one frame per context.");
        if (ctxs[i]->device != ctxs[0]->device) return fail(RM_ERR_INVALID, "A batch's contexts share one device.");
        int r = frame_check(ctxs[i], scenes[i], params[i], emitters ? emitters[i] : nullptr,
                            emitter_vec4s ? emitter_vec4s[i] : 0u, F[i]);
        if (r != RM_OK) return r;
        if (tree_of(ctxs[i], scenes[i]) != tree_of(ctxs[0], scenes[0]))
            return fail(RM_ERR_INVALID, "A batch's frames walk the same kind of tree (RM_PRECISION_QUANTIZED_BVH "
                                        "the same for all, and quantizable scenes).");
    }
    // Every context on the wavefront schedule (and rendering colour): the
    // jobs' wavefront passes, each stage one launch for all of them, replayed
    // from a hipGraph; otherwise one launch of the per-pixel pass.
    bool wavefront = true;
    for (uint32_t i = 0; i < n; ++i)
        wavefront = wavefront && ctxs[i]->options[RM_OPTION_SCHEDULE] == RM_SCHEDULE_WAVEFRONT && F[i].P.view == 0u;
    for (uint32_t i = 0; i < n; ++i) {
        int r = frame_alloc(ctxs[i], emitters ? emitters[i] : nullptr, emitter_vec4s ? emitter_vec4s[i] : 0u, F[i],
                            wavefront);
        if (r != RM_OK) return r;
    }
    // Light tracing: each frame with it enqueues its side stream's waits and
    // the clear of its photon counts; then one launch traces every such
    // frame's light paths, on the side stream of the first once the others'
    // side streams have reached it, and each frame's sort and gather follow
    // on its own side stream.
    for (uint32_t i = 0; i < n; ++i) {
        int r = frame_begin(ctxs[i], F[i], true);
        if (r != RM_OK) return r;
    }
    {
        rm::LightTable L;
        std::memset(&L, 0, sizeof L);
        rm::Features lu{};
        rm_context* lead = nullptr;
        uint32_t blocks = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!F[i].caustics) continue;
            if (!lead) {
                lead = ctxs[i];
                lu = F[i].f;
            }
            lu.textures |= F[i].f.textures;
            lu.dispersion |= F[i].f.dispersion;
            L.job[L.n] = light_job(ctxs[i], F[i]);
            L.begin[L.n++] = blocks;
            blocks += rm::light_blocks(F[i].photons);
        }
        if (lead) {
            L.begin[L.n] = blocks;
            for (uint32_t i = 0; i < n; ++i)
                if (F[i].caustics && ctxs[i] != lead) {
                    RM_HIP(hipEventRecord(ctxs[i]->side_ev, ctxs[i]->side), "ordering a batch's light tracing");
                    RM_HIP(hipStreamWaitEvent(lead->side, ctxs[i]->side_ev, 0), "ordering a batch's light tracing");
                }
            {
                Timed t(lead, RM_STAGE_LIGHT_TRACE, lead->side);
                rm::launch_light_trace_batch(L, lu, lead->side);
            }
            RM_HIP(hipGetLastError(), "launching the batch's light tracing");
            RM_HIP(hipEventRecord(lead->side_ev, lead->side), "ordering a batch's light tracing");
            for (uint32_t i = 0; i < n; ++i) {
                if (!F[i].caustics) continue;
                if (ctxs[i] != lead)
                    RM_HIP(hipStreamWaitEvent(ctxs[i]->side, lead->side_ev, 0), "ordering a batch's light tracing");
                int r = light_end(ctxs[i], F[i]);
                if (r != RM_OK) return r;
            }
        }
    }
    // One launch for all the path passes, on the first context's stream once
    // every context has queued what comes before; each context's stream then
    // waits for it and goes on with its own frame.
    rm::BatchTable T;
    std::memset(&T, 0, sizeof T);
    rm::Features u = F[0].f;
    uint32_t tiles = 0;
    for (uint32_t i = 0; i < n; ++i) {
        rm_context* c = ctxs[i];
        const uint32_t tx = (c->width + 15u) / 16u, ty = (c->height + 15u) / 16u;
        T.job[i] = rm::BatchJob{F[i].S, F[i].P, c->accum.as<float4>(), c->galb.as<float4>(), c->gnorm.as<float4>(),
                                F[i].comp, c->spill.as<uint32_t>(), F[i].px, tx};
        T.begin[i] = tiles;
        tiles += tx * ty;
        u.volumes |= F[i].f.volumes;
        u.textures |= F[i].f.textures;
        u.environment |= F[i].f.environment;
        u.dispersion |= F[i].f.dispersion;
    }
    T.begin[n] = tiles;
    T.n = n;
    rm_context* run = ctxs[0];
    for (uint32_t i = 1; i < n; ++i) {
        RM_HIP(hipEventRecord(ctxs[i]->batch_ev, ctxs[i]->stream), "ordering a batch");
        RM_HIP(hipStreamWaitEvent(run->stream, ctxs[i]->batch_ev, 0), "ordering a batch");
    }
    if (wavefront) {
        Timed t(run, RM_STAGE_TRACE);
        int r = wavefront_batch(run, ctxs, F, n, u);
        if (r != RM_OK) return r;
    } else {
        Timed t(run, RM_STAGE_TRACE);
        rm::launch_main_batch(T, u, run->stream);
    }
    RM_HIP(hipGetLastError(), "launching the batched trace");
    RM_HIP(hipEventRecord(run->batch_ev, run->stream), "ordering a batch");
    for (uint32_t i = 1; i < n; ++i)
        RM_HIP(hipStreamWaitEvent(ctxs[i]->stream, run->batch_ev, 0), "ordering a batch");
    // The denoiser: with two or more frames denoising (their guides of one
    // precision), each enqueues its resolve and prepare on its own stream,
    // then one launch per à-trous pass and one for compose run all of them on
    // the first one's post stream once every prepare is done; the others'
    // post streams wait for that before they record their frame denoised.
    uint32_t denoising = 0;
    bool fp16 = false, mixed = false;
    for (uint32_t i = 0; i < n; ++i) {
        if (!F[i].P.denoise) continue;
        const bool h = (ctxs[i]->options[RM_OPTION_PRECISION] & RM_PRECISION_FP16_GUIDES) != 0;
        mixed = mixed || (denoising > 0 && h != fp16);
        fp16 = h;
        ++denoising;
    }
    const bool batched_denoise = denoising >= 2 && !mixed;
    for (uint32_t i = 0; i < n; ++i) {
        int r = frame_end(ctxs[i], F[i], batched_denoise);
        if (r != RM_OK) return r;
    }
    if (batched_denoise) {
        rm::DenoiseTable D;
        std::memset(&D, 0, sizeof D);
        rm_context* lead = nullptr;
        uint32_t lead_frame = 0, blocks = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!F[i].P.denoise) continue;
            if (!lead) {
                lead = ctxs[i];
                lead_frame = i;
            }
            RM_HIP(hipStreamWaitEvent(lead->post, ctxs[i]->prepared_ev, 0), "ordering a batch's denoiser");
            D.job[D.n] = denoise_job(ctxs[i], F[i]);
            D.begin[D.n++] = blocks;
            blocks += rm::stencil_blocks(F[i].P.width, F[i].P.height);
        }
        D.begin[D.n] = blocks;
        {
            Timed t(lead, RM_STAGE_ATROUS, lead->post);
            rm::launch_denoise_batch(D, fp16, lead->post);
        }
        RM_HIP(hipGetLastError(), "launching the batch's denoiser");
        int r = denoise_recorded(lead, F[lead_frame]);
        if (r != RM_OK) return r;
        for (uint32_t i = 0; i < n; ++i) {
            if (!F[i].P.denoise || ctxs[i] == lead) continue;
            RM_HIP(hipStreamWaitEvent(ctxs[i]->post, lead->denoised_ev, 0), "ordering a batch's denoiser");
            r = denoise_recorded(ctxs[i], F[i]);
            if (r != RM_OK) return r;
        }
    }
    return RM_OK;
}

int rm_read(rm_context* c, int buffer, float* dst, size_t floats) {
    if (!c || !dst) return fail(RM_ERR_INVALID, "rm_read needs a context and a destination.");
    if (c->width == 0) return fail(RM_ERR_STATE, "Nothing has been rendered yet.");
    if (buffer < RM_BUFFER_ACCUM || buffer > RM_BUFFER_DENOISED)
        return fail(RM_ERR_INVALID, "Unknown buffer " + std::to_string(buffer) + ".");
    size_t have = (size_t)c->width * c->height * 4;
    if (floats > have) return fail(RM_ERR_INVALID, "rm_read asked for more than the image holds.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    if (c->denoise_pending) RM_HIP(hipStreamWaitEvent(c->stream, c->denoised_ev, 0), "waiting for the denoiser");
    const void* src = c->denoised.ptr;
    if (buffer != RM_BUFFER_DENOISED) {
        const float4* sums[3];
        int r = rmi::current_sums(c, c->stream, sums);
        if (r != RM_OK) return r;
        src = sums[buffer == RM_BUFFER_ACCUM ? 0 : buffer == RM_BUFFER_GUIDE_ALBEDO ? 1 : 2];
    }
    RM_HIP(hipMemcpyAsync(dst, src, floats * 4, hipMemcpyDeviceToHost, c->stream), "reading back pixels");
    RM_HIP(hipStreamSynchronize(c->stream), "reading back pixels");
    return RM_OK;
}

int rm_display(rm_context* c, uint8_t* rgba, size_t bytes) {
    if (!c || !rgba) return fail(RM_ERR_INVALID, "rm_display needs a context and a destination.");
    if (!c->rendered) return fail(RM_ERR_STATE, "Nothing has been rendered yet.");
    const rm::Params& P = c->last_params;
    const size_t need = (size_t)P.width * P.height * 4;
    if (P.width != c->width || P.height != c->height)
        return fail(RM_ERR_STATE, "The image was resized after the last frame.");
    if (bytes < need) return fail(RM_ERR_INVALID, "rm_display needs W·H·4 bytes.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    RM_HIP(c->display.ensure(need), "allocating the display image");
    if (c->denoise_pending) RM_HIP(hipStreamWaitEvent(c->stream, c->denoised_ev, 0), "waiting for the denoiser");
    const float4* sums[3];
    int r = rmi::current_sums(c, c->stream, sums);
    if (r != RM_OK) return r;
    rm::launch_display(P, sums[0], c->denoised.as<const float4>(), c->display.as<uchar4>(), c->stream);
    RM_HIP(hipGetLastError(), "launching the display pass");
    RM_HIP(hipMemcpyAsync(rgba, c->display.ptr, need, hipMemcpyDeviceToHost, c->stream), "reading back the display");
    RM_HIP(hipStreamSynchronize(c->stream), "waiting for the GPU");
    return RM_OK;
}

int rm_synchronize(rm_context* c) {
    if (!c) return fail(RM_ERR_INVALID, "rm_synchronize needs a context.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    if (c->denoise_pending) RM_HIP(hipStreamWaitEvent(c->stream, c->denoised_ev, 0), "waiting for the denoiser");
    RM_HIP(hipStreamSynchronize(c->stream), "waiting for the GPU");
    return RM_OK;
}

int rm_raycast(rm_context* c, const rm_scene* s, const float* rays, uint32_t n, int any_hit, uint32_t* hits,
               uint32_t* counters) {
    if (!c || !s || (!rays && n) || (!hits && n)) return fail(RM_ERR_INVALID, "rm_raycast needs rays and room for hits.");
    if (n == 0) return RM_OK;
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    RM_HIP(c->rays.ensure((size_t)n * 32), "allocating rays");
    RM_HIP(c->hits.ensure((size_t)n * 16), "allocating hits");
    if (counters) RM_HIP(c->counters.ensure((size_t)n * 8), "allocating counters");
    int r = ensure_spill(c, n > c->width * c->height ? n : c->width * c->height);
    if (r != RM_OK) return r;
    RM_HIP(hipMemcpyAsync(c->rays.ptr, rays, (size_t)n * 32, hipMemcpyHostToDevice, c->stream), "uploading rays");
    {
        Timed t(c, RM_STAGE_RAYCAST);
        rm::launch_raycast(dev_scene(c, s), c->rays.as<const float4>(), n, c->hits.as<uint4>(), any_hit != 0,
                           tree_of(c, s), c->spill.as<uint32_t>(), n, counters ? c->counters.as<uint2>() : nullptr, c->stream);
    }
    RM_HIP(hipGetLastError(), "launching the raycast");
    RM_HIP(hipMemcpyAsync(hits, c->hits.ptr, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream), "reading hits");
    if (counters)
        RM_HIP(hipMemcpyAsync(counters, c->counters.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream),
               "reading counters");
    RM_HIP(hipStreamSynchronize(c->stream), "tracing rays");
    return RM_OK;
}

int rm_raycast_static(rm_context* c, const rm_scene* s, const float* rays, uint32_t n, int any_hit, uint32_t* hits,
                      uint32_t* counters) {
    if (n > RM_RAYCAST_STATIC_MAX_RAYS)
        return fail(RM_ERR_INVALID, "rm_raycast_static traces at most " + std::to_string(RM_RAYCAST_STATIC_MAX_RAYS) +
                                        " rays a call (asked for " + std::to_string(n) + "); use rm_raycast, or split "
                                        "the rays.");
    if (!c || !s || (!rays && n) || (!hits && n))
        return fail(RM_ERR_INVALID, "rm_raycast_static needs a context, a scene, rays and room for hits.");
    if (n == 0) return RM_OK;
    if (!c->static_scratch) return fail(RM_ERR_STATE, "The context has no static raycast scratch.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    // The kernel's spill slot stride is n, so it indexes entries below
    // (STACK_LIMIT - TRACE_LDS_DEPTH) · n, inside the scratch.
    rm::launch_raycast(dev_scene(c, s), reinterpret_cast<const float4*>(rays), n, reinterpret_cast<uint4*>(hits),
                       any_hit != 0, tree_of(c, s), reinterpret_cast<uint32_t*>(c->static_scratch), n,
                       reinterpret_cast<uint2*>(counters), c->stream);
    RM_HIP(hipGetLastError(), "launching the raycast");
    return RM_OK;
}

int rm_timing(rm_context* c, int enabled) {
    if (!c) return fail(RM_ERR_INVALID, "rm_timing needs a context.");
    c->timing = enabled != 0;
    return RM_OK;
}

int rm_stage_times(rm_context* c, double* ms, uint32_t* launches, uint32_t n) {
    if (!c) return fail(RM_ERR_INVALID, "rm_stage_times needs a context.");
    RM_HIP(hipSetDevice(c->device), "selecting the device");
    int r = collect_times(c);
    if (r != RM_OK) return r;
    for (uint32_t s = 0; s < n && s < RM_STAGE_COUNT; ++s) {
        if (ms) ms[s] = c->timer.ms[s];
        if (launches) launches[s] = c->timer.launches[s];
    }
    return RM_OK;
}

int rm_reset_stage_times(rm_context* c) {
    if (!c) return fail(RM_ERR_INVALID, "rm_reset_stage_times needs a context.");
    int r = collect_times(c);
    if (r != RM_OK) return r;
    for (int s = 0; s < RM_STAGE_COUNT; ++s) {
        c->timer.ms[s] = 0.0;
        c->timer.launches[s] = 0;
    }
    return RM_OK;
}

int rm_tune(rm_context* c, const rm_scene* s, const void* params, const float* emitters, uint32_t emitter_vec4s,
            uint32_t frames, rm_tune_result* out) {
    if (!c || !s || !params || frames == 0 || frames > 64)
        return fail(RM_ERR_INVALID, "rm_tune needs a context, a scene, parameters and 1 to 64 frames.");
    rm::Params P;
    std::memcpy(&P, params, sizeof P);
    const uint32_t spp = P.samples_per_frame > 0 ? P.samples_per_frame : 1u;
    const uint32_t cus = (uint32_t)(c->compute_units > 0 ? c->compute_units : 1);
    // The candidates: schedule, persistent blocks, waves, LDS stack depth,
    // block shape, split depth.
    const int32_t schedules[RM_TUNE_CANDIDATES] = {RM_SCHEDULE_SPLIT,   RM_SCHEDULE_SPLIT,   RM_SCHEDULE_SPLIT,
                                                  RM_SCHEDULE_PIXEL,   RM_SCHEDULE_PIXEL,   RM_SCHEDULE_PIXEL,
                                                  RM_SCHEDULE_PIXEL,   RM_SCHEDULE_COMPACT, RM_SCHEDULE_SORTED,
                                                  RM_SCHEDULE_PERSISTENT, RM_SCHEDULE_PERSISTENT, RM_SCHEDULE_WAVEFRONT};
    const uint32_t blocks[RM_TUNE_CANDIDATES] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 2u * cus, 0u};
    const uint32_t waves[RM_TUNE_CANDIDATES] = {4u, 5u, 5u, 4u, 5u, 4u, 4u, 4u, 4u, 4u, 4u, 4u};
    const uint32_t depths[RM_TUNE_CANDIDATES] = {8u, 8u, 8u, 8u, 8u, 16u, 8u, 8u, 8u, 8u, 8u, 8u};
    const uint32_t shapes[RM_TUNE_CANDIDATES] = {RM_BLOCK_SQUARES, RM_BLOCK_SQUARES, RM_BLOCK_SQUARES, RM_BLOCK_SQUARES,
                                                RM_BLOCK_SQUARES, RM_BLOCK_SQUARES, RM_BLOCK_ROWS,    RM_BLOCK_SQUARES,
                                                RM_BLOCK_SQUARES, RM_BLOCK_SQUARES, RM_BLOCK_SQUARES, RM_BLOCK_SQUARES};
    const uint32_t splits[RM_TUNE_CANDIDATES] = {2u, 2u, 1u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    rm_tune_result res = {};
    constexpr int TUNED = 6;
    const int tuned[TUNED] = {RM_OPTION_SCHEDULE,    RM_OPTION_PERSISTENT_BLOCKS, RM_OPTION_WAVES,
                              RM_OPTION_STACK_DEPTH, RM_OPTION_BLOCK_SHAPE,       RM_OPTION_SPLIT_DEPTH};
    int64_t before[TUNED];
    for (int k = 0; k < TUNED; ++k) before[k] = c->options[tuned[k]];
    auto finish_all = [&]() -> int {
        RM_HIP(hipStreamSynchronize(c->stream), "tuning");
        RM_HIP(hipStreamSynchronize(c->side), "tuning");
        RM_HIP(hipStreamSynchronize(c->post), "tuning");
        return RM_OK;
    };
    double best_ms = 0.0;
    int r = RM_OK;
    const uint32_t mask = c->options[RM_OPTION_TUNE_CANDIDATES] != 0
                              ? (uint32_t)c->options[RM_OPTION_TUNE_CANDIDATES]
                              : (1u << RM_TUNE_CANDIDATES) - 1u;
    for (uint32_t k = 0; k < RM_TUNE_CANDIDATES; ++k) {
        res.schedule[k] = schedules[k];
        res.persistent_blocks[k] = blocks[k];
        res.waves[k] = waves[k];
        res.stack_depth[k] = depths[k];
        res.block_shape[k] = shapes[k];
        res.split_depth[k] = splits[k];
    }
    res.count = RM_TUNE_CANDIDATES;
    // Two rounds over the candidates the mask names, each keeping its faster
    // round: the first candidate of the first round otherwise pays for the
    // GPU coming out of idle. Between candidates the streams drain.
    for (uint32_t n = 0; n < 2 * RM_TUNE_CANDIDATES && r == RM_OK; ++n) {
        const uint32_t k = n % RM_TUNE_CANDIDATES;
        if (!(mask >> k & 1u)) continue;
        c->options[RM_OPTION_SCHEDULE] = schedules[k];
        c->options[RM_OPTION_PERSISTENT_BLOCKS] = blocks[k];
        c->options[RM_OPTION_WAVES] = waves[k];
        c->options[RM_OPTION_STACK_DEPTH] = depths[k];
        c->options[RM_OPTION_BLOCK_SHAPE] = shapes[k];
        c->options[RM_OPTION_SPLIT_DEPTH] = splits[k];
        // `frames` frames to warm up, the streams drained, then `frames`
        // more timed by events on the path stream. Light tracing, the gather
        // and the denoiser run beside the path pass on their own streams and
        // settle into a phase against it that the start decides and that
        // then holds: timed from a drained start, as a caller that
        // synchronises runs its frames, a candidate is timed in the phase it
        // will run in. (Timed straight on from the first frames, the sphere
        // frame at five waves ran in another phase, 7% faster than it ever
        // ran after a drain.)
        for (uint32_t f = 0; f < 2 * frames && r == RM_OK; ++f) {
            if (f == frames) {
                r = finish_all();
                if (r != RM_OK) break;
                RM_HIP(hipEventRecord(c->tune_ev[0], c->stream), "tuning");
            }
            P.sample_offset = f * spp;
            P.total_samples = (f + 1) * spp;
            r = rm_frame(c, s, &P, emitters, emitter_vec4s);
        }
        if (r == RM_OK) RM_HIP(hipEventRecord(c->tune_ev[1], c->stream), "tuning");
        if (r == RM_OK) r = finish_all();
        if (r != RM_OK) break;
        float ms = 0.0f;
        RM_HIP(hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]), "tuning");
        const double per_frame = (double)ms / frames;
        res.ms[k] = n < RM_TUNE_CANDIDATES ? per_frame : std::min(res.ms[k], per_frame);
    }
    // The candidates are in order of preference: a later one replaces the
    // best so far only when it is at least 1% faster, so that near-ties (the
    // sphere frame at four and five waves, 1.287 ms each over 16 frames, and
    // 1.27 against 1.32 ms over 64) go to the earlier.
    bool any = false;
    for (uint32_t k = 0; k < RM_TUNE_CANDIDATES; ++k)
        if ((mask >> k & 1u) && (!any || res.ms[k] < 0.99 * best_ms)) {
            best_ms = res.ms[k];
            res.best = k;
            any = true;
        }
    const uint32_t pick = res.best;
    const int64_t picked[TUNED] = {res.schedule[pick],     res.persistent_blocks[pick], res.waves[pick],
                                   res.stack_depth[pick], res.block_shape[pick],       res.split_depth[pick]};
    for (int k = 0; k < TUNED; ++k) c->options[tuned[k]] = r == RM_OK ? picked[k] : before[k];
    if (r != RM_OK) return r;
    if (out) *out = res;
    return RM_OK;
}

int rm_device_info(rm_context* c, char* name, size_t name_len, int* compute_units) {
    if (!c) return fail(RM_ERR_INVALID, "rm_device_info needs a context.");
    hipDeviceProp_t prop;
    RM_HIP(hipGetDeviceProperties(&prop, c->device), "reading device properties");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return RM_OK;
}

}  // extern "C"
