// The other schedules of the path pass (RM_OPTION_SCHEDULE): compacted
// (k_main_compact), persistent (k_main_persistent) and wavefront (k_wf_*).
// Each computes the same samples as k_main, bit for bit.
#include "path.hpp"

namespace rm {

// Items (pixels) a wave of the persistent path pass claims per atomic.
constexpr uint32_t CLAIM_ITEMS = 256u;

// The same samples with the block's live paths packed together between
// bounces. A block's 256 paths start together (coherent primary rays); after
// each bounce, when fewer waves could hold the paths still alive, their state
// moves through LDS to the first lanes of the block, so emptied waves wait at
// the barrier instead of running the bounce with a few live lanes. A path
// that ends leaves its sample in a per-lane staging record; once every path of
// the block has ended, each pixel's own thread adds its sample in order. Every
// path runs exactly the same arithmetic as in k_main, so the sums are
// bit-identical; only the colour view runs here. The barrier after every
// bounce ties a block's waves to its slowest one, so with short paths (1 spp
// per frame on the sphere scene) k_main is faster.
constexpr int COMPACT_LDS_DEPTH = 4;

// Sorting (SORT, RM_SCHEDULE_SORTED): each bounce is split in two at its
// closest hit, and while more than one wave's worth of paths is alive the
// block's live paths are reordered through the LDS slots twice a bounce:
// after tracing, by the kind of material hit (or none: the ray escaped), so
// a wave shades one kind of surface (diffuse, metal, plastic, glass, light)
// rather than branching between them; after shading, by the octant of the
// next ray's direction and whether the bounce was specular, so a wave
// traces rays heading the same way, which visit the tree's children in the
// same order. Lanes take slots by an LDS counting sort (a count per key,
// then each key's run in key order). Which lane a path lands on changes
// none of its arithmetic: its RNG is re-derived from its pixel and sample,
// and shading restarts the RNG's dimension from the bounce, so the sums
// stay bit-identical.
constexpr uint32_t SORT_KEYS = 16;
RM_D uint32_t direction_key(const PathState& ps) {
    return (ps.d.x < 0.0f ? 1u : 0u) | (ps.d.y < 0.0f ? 2u : 0u) | (ps.d.z < 0.0f ? 4u : 0u) |
           (ps.last_delta ? 8u : 0u);
}
RM_D uint32_t material_key(const DevScene& S, const Hit& h) {
    if (h.prim == NONE) return 7u;
    const uint32_t m = __float_as_uint(S.prims[(size_t)h.prim * 4].y);
    return __float_as_uint(S.mats[(size_t)m * 6].w) & 7u;
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, bool SORT>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main_compact(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots,
    float4* __restrict__ staging) {
    __shared__ uint2 lds[COMPACT_LDS_DEPTH * BLOCK_THREADS];
    __shared__ uint32_t slot_words[SLOT_WORDS * BLOCK_THREADS];
    __shared__ uint32_t wave_live[BLOCK_THREADS / 64];
    __shared__ uint32_t key_count[SORT_KEYS];
    __shared__ uint4 hit_slots[SORT ? BLOCK_THREADS : 1];   // a sorted path's closest hit rides with it
    const Ctx C{S, &P};
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    const uint32_t gid = (blockIdx.y * gridDim.x + blockIdx.x) * BLOCK_THREADS + t;
    uint32_t x, y;
    tile_pixel(x, y);
    const bool valid = x < P.width && y < P.height;
    const uint32_t index = valid ? y * P.width + x : 0u;
    auto stack = BlockStack<COMPACT_LDS_DEPTH>::make(lds, spill, gid, slots);
    SlotStore st{slot_words, BLOCK_THREADS};
    float4* const stage = staging + (size_t)(gid - t) * 3;  // this block's 256 records of 3 float4
    PixelSums sums;
    sums.clear();
    for (uint32_t k = 0; k < P.samples_per_frame; k++) {
        const uint32_t sample = P.sample_offset + k;
        Rng g = {};
        PathState ps = {};
        Guides gd = {};
        uint32_t pix = t;
        bool live = valid;
        if (live) path_begin(P, g, ps, gd, x, y, index, sample);
        // The number of live paths in the block (and in the waves before this one).
        auto count_live = [&](uint64_t& mask, uint32_t& before, uint32_t& nonempty) {
            mask = __ballot(live);
            if (lane == 0) wave_live[wave] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t total = 0;
            before = nonempty = 0;
            for (uint32_t w = 0; w < BLOCK_THREADS / 64; w++) {
                const uint32_t n = wave_live[w];
                before += w < wave ? n : 0u;
                total += n;
                nonempty += n > 0u ? 1u : 0u;
            }
            return total;
        };
        // Moves the live paths (and their hits) to slots 0 .. total-1 in key order.
        auto regroup = [&](uint32_t key, uint32_t total, Hit& hit) {
            if (t < SORT_KEYS) key_count[t] = 0u;
            __syncthreads();
            const uint32_t rank = live ? atomicAdd(&key_count[key], 1u) : 0u;
            __syncthreads();
            uint32_t start = 0;
            for (uint32_t k2 = 0; k2 < SORT_KEYS; k2++) start += k2 < key ? key_count[k2] : 0u;
            if (live) {
                slot_save(st, start + rank, ps, gd, pix);
                hit_slots[SORT ? start + rank : 0] =
                    make_uint4(__float_as_uint(hit.t), hit.prim, __float_as_uint(hit.u), __float_as_uint(hit.v));
            }
            __syncthreads();
            live = t < total;
            if (live) {
                slot_load(st, t, P, sample, ps, gd, g, pix);
                const uint4 h = hit_slots[SORT ? t : 0];
                hit = {__uint_as_float(h.x), h.y, __uint_as_float(h.z), __uint_as_float(h.w)};
            }
            __syncthreads();
        };
        if constexpr (SORT) {
            for (;;) {
                uint64_t mask;
                uint32_t before, nonempty;
                NoCount nc;
                Hit hit = {INF, NONE, 0.0f, 0.0f};
                if (live) hit = trace<false>(C.S, ps.o, ps.d, INF, stack, nc);
                uint32_t total = count_live(mask, before, nonempty);
                if (total == 0u) break;
                // (Either way a barrier: no wave rewrites wave_live while another still reads it.)
                if (total > 64u) regroup(live ? material_key(C.S, hit) : 0u, total, hit);
                else __syncthreads();
                if (live) {
                    PendingNee pn;
                    const bool ended = path_shade<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, ps, gd, hit, pn);
                    if (pn.active && trace<true>(C.S, pn.o, pn.d, pn.reach, stack, nc).prim == NONE)
                        apply_nee<VOLUMES>(C, ps, pn);
                    if (ended) {
                        stage[pix * 3] = make_float4(ps.result.x, ps.result.y, ps.result.z, 0.0f);
                        stage[pix * 3 + 1] = make_float4(gd.albedo.x, gd.albedo.y, gd.albedo.z, 0.0f);
                        stage[pix * 3 + 2] = gd.normal;
                        live = false;
                    }
                }
                total = count_live(mask, before, nonempty);
                if (total == 0u) break;
                if (total > 64u) regroup(live ? direction_key(ps) : 0u, total, hit);
                else __syncthreads();
            }
        } else
            for (;;) {
            if (live && path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
                stage[pix * 3] = make_float4(ps.result.x, ps.result.y, ps.result.z, 0.0f);
                stage[pix * 3 + 1] = make_float4(gd.albedo.x, gd.albedo.y, gd.albedo.z, 0.0f);
                stage[pix * 3 + 2] = gd.normal;
                live = false;
            }
            const uint64_t mask = __ballot(live);
            if (lane == 0) wave_live[wave] = (uint32_t)__popcll(mask);
            __syncthreads();
            uint32_t total = 0, before = 0, nonempty = 0;
            for (uint32_t w = 0; w < BLOCK_THREADS / 64; w++) {
                uint32_t n = wave_live[w];
                before += w < wave ? n : 0u;
                total += n;
                nonempty += n > 0u ? 1u : 0u;
            }
            if (total == 0u) break;
            if ((total + 63u) / 64u < nonempty) {
                // Pack the live paths into lanes 0 .. total-1, in lane order.
                if (live) slot_save(st, before + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)), ps, gd, pix);
                __syncthreads();
                live = t < total;
                if (live) slot_load(st, t, P, sample, ps, gd, g, pix);
            }
            __syncthreads();
        }
        __syncthreads();
        if (valid) {
            float4 r = stage[t * 3], a = stage[t * 3 + 1];
            Guides own;
            own.albedo = {a.x, a.y, a.z};
            own.normal = stage[t * 3 + 2];
            sums.add(P, v3(r.x, r.y, r.z), own, true);
        }
        __syncthreads();
    }
    if (valid) sums.store(P, index, accum, guide_albedo, guide_normal, comp);
}

// The same samples with lanes kept busy: waves claim 8×8 pixel tiles from a
// global counter (one atomic per CLAIM_ITEMS pixels), and a lane whose path
// ends starts its pixel's next sample, or the next unclaimed pixel, while the
// rest of the wave keeps bouncing. A pixel's samples stay in one lane, in
// order, so every sum is bit-identical to k_main's; only the colour view runs
// here. Regeneration mixes fresh primary rays with deep secondary ones in a
// wave, which costs the primary rays' coherence: at one sample per pixel per
// frame, k_main is the faster schedule.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) void k_main_persistent(DevScene S, Params P, float4* __restrict__ accum,
                                                                   float4* __restrict__ guide_albedo,
                                                                   float4* __restrict__ guide_normal,
                                                                   float4* __restrict__ comp, uint32_t* spill,
                                                                   uint32_t slots, uint32_t* counter) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t gid = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    auto stack = BlockStack<MAIN_LDS_DEPTH>::make(lds, spill, gid, slots);
    const uint32_t lane = __lane_id();
    const uint32_t tiles_x = (P.width + 7u) / 8u;
    const uint32_t items = tiles_x * ((P.height + 7u) / 8u) * 64u;

    bool alive = false;
    uint32_t x = 0, y = 0, index = 0, k = 0;
    Rng g = {};
    PathState ps = {};
    Guides gd = {};
    PixelSums sums;
    sums.clear();
    uint32_t next = 0, end = 0;  // the wave's claimed, unassigned items
    bool exhausted = false;

    for (;;) {
        uint64_t need = __ballot(!alive);
        while (need != 0 && !exhausted) {
            if (next >= end) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(counter, CLAIM_ITEMS);
                base = __shfl(base, 0);
                if (base >= items) {
                    exhausted = true;
                    break;
                }
                next = base;
                end = min(base + CLAIM_ITEMS, items);
            }
            uint32_t avail = end - next;
            uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            if (((need >> lane) & 1ull) && rank < avail) {
                uint32_t item = next + rank, tile = item >> 6, l = item & 63u;
                x = (tile % tiles_x) * 8u + (l & 7u);
                y = (tile / tiles_x) * 8u + (l >> 3);
                if (x < P.width && y < P.height) {
                    alive = true;
                    index = y * P.width + x;
                    k = 0;
                    sums.clear();
                    path_begin(P, g, ps, gd, x, y, index, P.sample_offset);
                }
            }
            next += min((uint32_t)__popcll(need), avail);
            need = __ballot(!alive);
        }
        if (__ballot(alive) == 0) break;
        if (alive && path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
            sums.add(P, ps.result, gd, true);
            if (++k < P.samples_per_frame) {
                path_begin(P, g, ps, gd, x, y, index, P.sample_offset + k);
            } else {
                sums.store(P, index, accum, guide_albedo, guide_normal, comp);
                alive = false;
            }
        }
    }
}

// ------------------------------------------------------------------ wavefront
// RM_SCHEDULE_WAVEFRONT: a frame's camera paths advanced one bounce at a time
// for every pixel, each stage a kernel of its own, a thread per queue entry:
//
//   begin   path_begin for every pixel's sample: the first live queue;
//   extend  the closest hit of each live path (a lean traversal kernel);
//   shade   path_shade at that hit; the path goes on to the next live queue
//           or, ended, to the done queue, its shadow ray to the shadow queue;
//   shadow  the any-hit test of each shadow ray, then apply_nee to its
//           path's result wherever that path went;
//   finish  PixelSums::add for the paths that ended;
//   store   PixelSums::store once the frame's samples are all in.
//
// A path's state lives at its queue slot, double-buffered between bounces,
// so every stage reads and writes its queue in order; slots are handed out
// one atomic per wave, so a wave's writes are contiguous. The same functions
// in the same order per path as k_main, each pixel's samples added in sample
// order: the accumulation is the same, bit for bit.
constexpr int NEE_WORDS = 14;    // o, d, vec, reach, phase, mis, pdf, medium
constexpr int DONE_WORDS = 11;   // result, guide albedo, guide normal, pixel
static_assert(WF_STATE_WORDS == SLOT_WORDS && WF_NEE_WORDS == NEE_WORDS + 1 && WF_DONE_WORDS == DONE_WORDS,
              "the host sizes these buffers");

RM_D SlotStore wf_state(const Wavefront& W, uint32_t which) { return SlotStore{W.state[which], W.n}; }

RM_D PixelSums wf_sums(const Wavefront& W, uint32_t i) {
    const float4 a = W.sums[i], b = W.sums[(size_t)W.n + i], c = W.sums[(size_t)2 * W.n + i];
    PixelSums s;
    s.c = {a.x, a.y, a.z};
    s.w = a.w;
    s.albedo = {b.x, b.y, b.z};
    s.normal = c;
    return s;
}

RM_D void wf_put_sums(const Wavefront& W, uint32_t i, const PixelSums& s) {
    W.sums[i] = make_float4(s.c.x, s.c.y, s.c.z, s.w);
    W.sums[(size_t)W.n + i] = make_float4(s.albedo.x, s.albedo.y, s.albedo.z, 0.0f);
    W.sums[(size_t)2 * W.n + i] = s.normal;
}

// The stages' work for one queue entry i, shared by the one-frame kernels
// below and the cross-job kernels (k_wfb_*) after them.

// Every pixel's sample `sample`: its camera path, in pixel order in live
// state 0; the first sample of the frame also clears the pixel's sums.
RM_D void wf_begin_one(const Params& P, const Wavefront& W, uint32_t i, uint32_t sample, bool first) {
    Rng g;
    PathState ps;
    Guides gd;
    path_begin(P, g, ps, gd, i % P.width, i / P.width, i, sample);
    SlotStore st = wf_state(W, 0);
    slot_save(st, i, ps, gd, i);
    if (first) {
        PixelSums s;
        s.clear();
        wf_put_sums(W, i, s);
    }
}

template <class STACK>
RM_D void wf_extend_one(const DevScene& S, const Wavefront& W, uint32_t cur, uint32_t i, STACK& stack) {
    const SlotStore st = wf_state(W, cur);
    NoCount nc;
    const Hit h = trace<false>(S, st.get3(i, 0), st.get3(i, 3), INF, stack, nc);
    W.hits[i] = make_uint4(__float_as_uint(h.t), h.prim, __float_as_uint(h.u), __float_as_uint(h.v));
}

// Shades entry i when live (i below the live count); every lane of the wave
// calls it, live or not, so the slot ballots see the whole wave.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
RM_D void wf_shade_one(const Ctx& C, const Wavefront& W, uint32_t cur, uint32_t sample, uint32_t i, bool live) {
    const uint32_t next = cur ^ 1u;
    bool ended = false;
    PendingNee pn;
    pn.active = false;
    Rng g;
    PathState ps;
    Guides gd;
    uint32_t pix = 0;
    if (live) {
        const SlotStore st = wf_state(W, cur);
        slot_load_wf(st, i, sample, ps, gd, g, pix);
        const uint4 h = W.hits[i];
        const Hit hit = {__uint_as_float(h.x), h.y, __uint_as_float(h.z), __uint_as_float(h.w)};
        ended = path_shade<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, ps, gd, hit, pn);
    }
    const uint32_t go = wave_slot(W.counts + next, live && !ended);
    const uint32_t stop = wave_slot(W.counts + 3, live && ended);
    const uint32_t shadow = wave_slot(W.counts + 2, live && pn.active);
    if (!live) return;
    if (ended) {
        SlotStore dn{W.done, W.n};
        dn.put3(stop, 0, ps.result);
        dn.put3(stop, 3, gd.albedo);
        dn.put(stop, 6, gd.normal.x);
        dn.put(stop, 7, gd.normal.y);
        dn.put(stop, 8, gd.normal.z);
        dn.put(stop, 9, gd.normal.w);
        dn.putu(stop, 10, pix);
    } else {
        SlotStore st = wf_state(W, next);
        slot_save(st, go, ps, gd, pix);
    }
    if (pn.active) {
        SlotStore nz{W.nee, W.n};
        nz.put3(shadow, 0, pn.o);
        nz.put3(shadow, 3, pn.d);
        nz.put3(shadow, 6, pn.vec);
        nz.put(shadow, 9, pn.reach);
        nz.put(shadow, 10, pn.phase);
        nz.put(shadow, 11, pn.mis);
        nz.put(shadow, 12, pn.pdf);
        nz.putu(shadow, 13, pn.medium ? 1u : 0u);
        // Where the path's result went: its done slot (top bit) or its next live slot.
        nz.putu(shadow, 14, ended ? (stop | 0x80000000u) : go);
    }
}

template <bool VOLUMES, class STACK>
RM_D void wf_shadow_one(const Ctx& C, const Wavefront& W, uint32_t next, uint32_t i, STACK& stack) {
    const SlotStore nz{W.nee, W.n};
    PendingNee pn;
    pn.o = nz.get3(i, 0);
    pn.d = nz.get3(i, 3);
    pn.vec = nz.get3(i, 6);
    pn.reach = nz.get(i, 9);
    pn.phase = nz.get(i, 10);
    pn.mis = nz.get(i, 11);
    pn.pdf = nz.get(i, 12);
    pn.medium = nz.getu(i, 13) != 0u;
    pn.active = true;
    NoCount nc;
    if (trace<true>(C.S, pn.o, pn.d, pn.reach, stack, nc).prim == NONE) {
        const uint32_t to = nz.getu(i, 14);
        SlotStore dst = (to & 0x80000000u) ? SlotStore{W.done, W.n} : wf_state(W, next);
        const uint32_t slot = to & 0x7fffffffu;
        const int word = (to & 0x80000000u) ? 0 : 9;   // the result's first word
        PathState ps = {};
        ps.result = dst.get3(slot, word);
        apply_nee<VOLUMES>(C, ps, pn);
        dst.put3(slot, word, ps.result);
    }
}

RM_D void wf_finish_one(const Params& P, const Wavefront& W, uint32_t i) {
    const SlotStore dn{W.done, W.n};
    Guides gd;
    gd.albedo = dn.get3(i, 3);
    gd.normal = make_float4(dn.get(i, 6), dn.get(i, 7), dn.get(i, 8), dn.get(i, 9));
    const uint32_t pix = dn.getu(i, 10);
    PixelSums s = wf_sums(W, pix);
    s.add(P, dn.get3(i, 0), gd, true);
    wf_put_sums(W, pix, s);
}

// One thread per queue entry: every stage's grid covers all n paths, and the
// threads past the queue's length (read on the device) return at once, so
// finished waves give way to new ones as in a kernel launched per ray.
#define WF_EACH(i, count) if (const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x; i < (count))

__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_begin(Params P, Wavefront W, uint32_t sample, uint32_t first) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        W.counts[0] = W.n;
        W.counts[1] = W.counts[2] = W.counts[3] = 0u;
    }
    WF_EACH(i, W.n) wf_begin_one(P, W, i, sample, first != 0u);
}

template <int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_extend(DevScene S, Wavefront W, uint32_t cur, uint32_t* spill,
                                                             uint32_t slots) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    // Nothing reads these until the shade kernel after this one.
    if (blockIdx.x == 0 && threadIdx.x == 0) W.counts[cur ^ 1u] = W.counts[2] = W.counts[3] = 0u;
    WF_EACH(i, W.counts[cur]) {
        auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, spill, i, slots);
        wf_extend_one(S, W, cur, i, stack);
    }
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_shade(DevScene S, Params P, Wavefront W, uint32_t cur,
                                                            uint32_t sample) {
    const Ctx C{S, &P};
    const uint32_t count = W.counts[cur];
    // Whole waves reach the slot atomics together, so their ballots see every lane.
    if (blockIdx.x * BLOCK_THREADS >= count) return;
    const uint32_t i = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    wf_shade_one<VOLUMES, TEXTURES, ENV, DISPERSION>(C, W, cur, sample, i, i < count);
}

template <bool VOLUMES, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_shadow(DevScene S, Params P, Wavefront W, uint32_t next,
                                                             uint32_t* spill, uint32_t slots) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    WF_EACH(i, W.counts[2]) {
        auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, spill, i, slots);
        wf_shadow_one<VOLUMES>(C, W, next, i, stack);
    }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_finish(Params P, Wavefront W) {
    WF_EACH(i, W.counts[3]) wf_finish_one(P, W, i);
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wf_store(Params P, Wavefront W, float4* __restrict__ accum,
                                                            float4* __restrict__ galb, float4* __restrict__ gnorm,
                                                            float4* __restrict__ comp) {
    WF_EACH(i, W.n) wf_sums(W, i).store(P, i, accum, galb, gnorm, comp);
}

// ------------------------------------------------------- cross-job wavefront
// Several jobs' wavefront path passes (rm_frame_batch, every context on
// RM_SCHEDULE_WAVEFRONT), each stage one launch for all of them. The job
// table lives in device memory, so the launches' arguments don't change from
// frame to frame and the whole sequence is captured once in a hipGraph and
// replayed. The grids are persistent: a fixed number of blocks takes the
// jobs' 256-entry chunks in turn (job j's are [begin[j], begin[j + 1])); a
// chunk lies within one job, so its job and Params are uniform over the
// block, read by scalar loads. Each job keeps its own queues and counters, so
// its paths are shaded in the same order, with the same functions, as by the
// one-job kernels above: the images are the same, bit for bit.
RM_D uint32_t wfb_job(const WfTable& T, uint32_t chunk) {
    uint32_t j = 0;
    while (j + 1u < T.n && chunk >= T.begin[j + 1u]) ++j;
    return j;
}

#define WFB_EACH_CHUNK(T, J, base)                                                        \
    for (uint32_t chunk_ = blockIdx.x; chunk_ < (T).begin[(T).n]; chunk_ += gridDim.x)    \
        if (const uint32_t j_ = wfb_job((T), chunk_), base = (chunk_ - (T).begin[j_]) * BLOCK_THREADS; true) \
            if (const WfJob& J = (T).job[j_]; true)

__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_begin(const WfTable* __restrict__ table, uint32_t k) {
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        const bool on = k < J.P.samples_per_frame;
        if (base == 0u && threadIdx.x == 0) {
            J.W.counts[0] = on ? J.W.n : 0u;
            J.W.counts[1] = J.W.counts[2] = J.W.counts[3] = 0u;
        }
        const uint32_t i = base + threadIdx.x;
        if (on && i < J.W.n) wf_begin_one(J.P, J.W, i, J.P.sample_offset + k, k == 0u);
    }
}

template <int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_extend(const WfTable* __restrict__ table, uint32_t cur) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        if (base == 0u && threadIdx.x == 0) J.W.counts[cur ^ 1u] = J.W.counts[2] = J.W.counts[3] = 0u;
        const uint32_t i = base + threadIdx.x;
        if (i < J.W.counts[cur]) {
            auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, J.spill, i, J.slots);
            wf_extend_one(J.S, J.W, cur, i, stack);
        }
    }
}

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_shade(const WfTable* __restrict__ table, uint32_t cur,
                                                             uint32_t k) {
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        const uint32_t count = J.W.counts[cur];
        if (base >= count) continue;   // the whole block: whole waves reach the slot ballots
        const Ctx C{J.S, &J.P};
        const uint32_t i = base + threadIdx.x;
        wf_shade_one<VOLUMES, TEXTURES, ENV, DISPERSION>(C, J.W, cur, J.P.sample_offset + k, i, i < count);
    }
}

template <bool VOLUMES, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_shadow(const WfTable* __restrict__ table, uint32_t next) {
    __shared__ uint2 lds[TRACE_LDS_DEPTH * BLOCK_THREADS];
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        const uint32_t i = base + threadIdx.x;
        if (i < J.W.counts[2]) {
            const Ctx C{J.S, &J.P};
            auto stack = BlockStack<TRACE_LDS_DEPTH, TREE>::make(lds, J.spill, i, J.slots);
            wf_shadow_one<VOLUMES>(C, J.W, next, i, stack);
        }
    }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_finish(const WfTable* __restrict__ table) {
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        const uint32_t i = base + threadIdx.x;
        if (i < J.W.counts[3]) wf_finish_one(J.P, J.W, i);
    }
}

__global__ __launch_bounds__(BLOCK_THREADS) void k_wfb_store(const WfTable* __restrict__ table) {
    const WfTable& T = *table;
    WFB_EACH_CHUNK(T, J, base) {
        const uint32_t i = base + threadIdx.x;
        if (i < J.W.n) wf_sums(J.W, i).store(J.P, i, J.accum, J.galb, J.gnorm, J.comp);
    }
}

#undef WFB_EACH_CHUNK
#undef WF_EACH

// ------------------------------------------------------------------ launchers
template <bool V, bool T, bool E, bool D>
static void launch_persistent_v(const LaunchMain& L, hipStream_t stream) {
    hipLaunchKernelGGL((k_main_persistent<V, T, E, D>), dim3(L.persistent_blocks), dim3(BLOCK_THREADS), 0, stream,
                       L.S, L.P, L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots, L.counter);
}

void launch_main_persistent(const LaunchMain& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_persistent_v<false, false, false, false>(L, stream); break;
    case 1: launch_persistent_v<false, false, false, true>(L, stream); break;
    case 2: launch_persistent_v<false, false, true, false>(L, stream); break;
    case 3: launch_persistent_v<false, false, true, true>(L, stream); break;
    case 8: launch_persistent_v<true, false, false, false>(L, stream); break;
    case 10: launch_persistent_v<true, false, true, false>(L, stream); break;
    default: launch_persistent_v<true, true, true, true>(L, stream); break;
    }
}

template <bool V, bool T, bool E, bool D>
static void launch_compact_v(const LaunchMain& L, hipStream_t stream) {
    dim3 grid((L.P.width + 15) / 16, (L.P.height + 15) / 16);
    if (L.sorted)
        hipLaunchKernelGGL((k_main_compact<V, T, E, D, true>), grid, dim3(BLOCK_THREADS), 0, stream, L.S, L.P, L.accum,
                           L.galb, L.gnorm, L.comp, L.spill, L.slots, L.staging);
    else
        hipLaunchKernelGGL((k_main_compact<V, T, E, D, false>), grid, dim3(BLOCK_THREADS), 0, stream, L.S, L.P,
                           L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots, L.staging);
}

void launch_main_compact(const LaunchMain& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_compact_v<false, false, false, false>(L, stream); break;
    case 1: launch_compact_v<false, false, false, true>(L, stream); break;
    case 2: launch_compact_v<false, false, true, false>(L, stream); break;
    case 3: launch_compact_v<false, false, true, true>(L, stream); break;
    case 8: launch_compact_v<true, false, false, false>(L, stream); break;
    case 10: launch_compact_v<true, false, true, false>(L, stream); break;
    default: launch_compact_v<true, true, true, true>(L, stream); break;
    }
}

template <bool V, bool T, bool E, bool D>
static int resident_blocks_v() {
    static int blocks = -1;
    if (blocks < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_main_persistent<V, T, E, D>, BLOCK_THREADS, 0) !=
            hipSuccess)
            n = 1;
        blocks = n > 0 ? n : 1;
    }
    return blocks;
}

int main_resident_blocks_per_cu(const Features& f) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: return resident_blocks_v<false, false, false, false>();
    case 1: return resident_blocks_v<false, false, false, true>();
    case 2: return resident_blocks_v<false, false, true, false>();
    case 3: return resident_blocks_v<false, false, true, true>();
    case 8: return resident_blocks_v<true, false, false, false>();
    case 10: return resident_blocks_v<true, false, true, false>();
    default: return resident_blocks_v<true, true, true, true>();
    }
}

template <bool V, bool T, bool E, bool D>
static void launch_wavefront_v(const LaunchWavefront& L, hipStream_t stream) {
    const dim3 grid(L.blocks), block(BLOCK_THREADS);
    for (uint32_t k = 0; k < L.P.samples_per_frame; ++k) {
        const uint32_t sample = L.P.sample_offset + k;
        hipLaunchKernelGGL(k_wf_begin, grid, block, 0, stream, L.P, L.W, sample, k == 0 ? 1u : 0u);
        // Every path ends within max_bounces + 1 bounces.
        for (uint32_t b = 0; b <= L.P.max_bounces; ++b) {
            const uint32_t cur = b & 1u;
            with_tree(L.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wf_extend<decltype(t)::value>), grid, block, 0, stream, L.S, L.W, cur, L.spill,
                                   L.slots);
            });
            hipLaunchKernelGGL((k_wf_shade<V, T, E, D>), grid, block, 0, stream, L.S, L.P, L.W, cur, sample);
            with_tree(L.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wf_shadow<V, decltype(t)::value>), grid, block, 0, stream, L.S, L.P, L.W,
                                   cur ^ 1u, L.spill, L.slots);
            });
            hipLaunchKernelGGL(k_wf_finish, grid, block, 0, stream, L.P, L.W);
        }
    }
    hipLaunchKernelGGL(k_wf_store, grid, block, 0, stream, L.P, L.W, L.accum, L.galb, L.gnorm, L.comp);
}

template <bool V, bool T, bool E, bool D>
static void launch_wavefront_batch_v(const WfTable* table, const WfBatchShape& B, hipStream_t stream) {
    const dim3 grid(B.blocks), block(BLOCK_THREADS);
    for (uint32_t k = 0; k < B.samples; ++k) {
        hipLaunchKernelGGL(k_wfb_begin, grid, block, 0, stream, table, k);
        for (uint32_t b = 0; b < B.bounces; ++b) {
            const uint32_t cur = b & 1u;
            with_tree(B.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wfb_extend<decltype(t)::value>), grid, block, 0, stream, table, cur);
            });
            hipLaunchKernelGGL((k_wfb_shade<V, T, E, D>), grid, block, 0, stream, table, cur, k);
            with_tree(B.tree, [&](auto t) {
                hipLaunchKernelGGL((k_wfb_shadow<V, decltype(t)::value>), grid, block, 0, stream, table, cur ^ 1u);
            });
            hipLaunchKernelGGL(k_wfb_finish, grid, block, 0, stream, table);
        }
    }
    hipLaunchKernelGGL(k_wfb_store, grid, block, 0, stream, table);
}

void launch_wavefront_batch(const WfTable* table, const WfBatchShape& B, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_wavefront_batch_v<false, false, false, false>(table, B, stream); break;
    case 1: launch_wavefront_batch_v<false, false, false, true>(table, B, stream); break;
    case 2: launch_wavefront_batch_v<false, false, true, false>(table, B, stream); break;
    case 3: launch_wavefront_batch_v<false, false, true, true>(table, B, stream); break;
    case 8: launch_wavefront_batch_v<true, false, false, false>(table, B, stream); break;
    case 10: launch_wavefront_batch_v<true, false, true, false>(table, B, stream); break;
    default: launch_wavefront_batch_v<true, true, true, true>(table, B, stream); break;
    }
}

void launch_wavefront(const LaunchWavefront& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_wavefront_v<false, false, false, false>(L, stream); break;
    case 1: launch_wavefront_v<false, false, false, true>(L, stream); break;
    case 2: launch_wavefront_v<false, false, true, false>(L, stream); break;
    case 3: launch_wavefront_v<false, false, true, true>(L, stream); break;
    case 8: launch_wavefront_v<true, false, false, false>(L, stream); break;
    case 10: launch_wavefront_v<true, false, true, false>(L, stream); break;
    default: launch_wavefront_v<true, true, true, true>(L, stream); break;
    }
}

}  // namespace rm
