// RM_SCHEDULE_SPLIT: the path pass in two kernels, split after a path's
// first bounces. Each returns the same samples as k_main, bit for bit.
//
//   head  one thread per pixel, as k_main: a 16×16 block of 8×8 wave squares
//         traces the coherent primary rays and the first `depth` bounces;
//         paths that end there add their sample to the pixel at once, the
//         rest are queued (27 words each, one atomic per wave);
//   tail  persistent waves run the queued paths to their ends; a lane whose
//         path ends takes the next queued one, so the waves stay full of
//         live paths until the queue runs dry.
//
// A wave of k_main runs as long as its longest path, and on the 1080p sphere
// frame 38% of primary rays escape at once and only a quarter of the paths
// live past their second bounce, so most of a wave's lanes idle through its
// last bounces. Here the long paths of many waves share the tail's waves.
// The split costs one store and one load of a survivor's state (108 bytes
// each way). Only a frame of one sample per pixel in the colour view is
// split (k_main runs the others): then each pixel has exactly one path, and
// whichever kernel ends it adds the pixel's only sample, so the order of
// adds is k_main's. The RNG of a queued path is re-derived from its pixel
// and sample, and shading restarts its dimension from the bounce.
//
// Both kernels are compiled for WAVES waves per SIMD (RM_OPTION_WAVES, as for
// k_main): split.hip instantiates them at four, split5.hip at five.
#pragma once
#include "main.hpp"

namespace rm {

template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE, int WAVES>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_split_head(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots, SlotStore queue,
    uint32_t* counts, uint32_t depth) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    tile_pixel(x, y);
    const bool valid = x < P.width && y < P.height;
    const Ctx C{S, &P};
    const uint32_t index = valid ? y * P.width + x : 0u;
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, index, slots);
    Rng g;
    PathState ps;
    Guides gd;
    bool ended = true;
    if (valid) {
        path_begin(P, g, ps, gd, x, y, index, P.sample_offset);
        for (;;) {
            ended = path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd);
            if (ended || ps.depth >= depth) break;
        }
    }
    // Every lane of the wave takes part in the slot ballot.
    const uint32_t slot = wave_slot(counts, valid && !ended);
    if (!valid) return;
    if (ended) {
        PixelSums sums;
        sums.clear();
        sums.add(P, ps.result, gd, true);
        sums.store(P, index, accum, guide_albedo, guide_normal, comp);
    } else {
        slot_save(queue, slot, ps, gd, index);
    }
}

// counts[0]: paths the head queued; counts[1]: the tail's claim counter.
// A lane whose path ends takes the next queued path while the rest of its
// wave goes on bouncing, so a wave's lanes stay busy until the queue runs
// dry: the queued paths are all past their first bounces, with no coherent
// primary rays left to keep together. The lanes that need a path claim
// their slots with one atomic a wave.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE, int WAVES>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_split_tail(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots, SlotStore queue,
    uint32_t* counts) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    const Ctx C{S, &P};
    const uint32_t gid = blockIdx.x * BLOCK_THREADS + threadIdx.x;
    auto stack = BlockStack<MAIN_LDS_DEPTH, TREE>::make(lds, spill, gid, slots);
    const uint32_t queued = counts[0];
    const uint32_t lane = __lane_id();
    Rng g = {};
    PathState ps = {};
    Guides gd = {};
    uint32_t index = 0;
    bool alive = false, drained = false;   // drained is the same on every lane of the wave
    for (;;) {
        const uint64_t need = __ballot(!alive);
        if (need != 0ull && !drained) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&counts[1], (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            drained = base + (uint32_t)__popcll(need) >= queued;
            if (!alive) {
                const uint32_t i = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
                if (i < queued) {
                    slot_load_wf(queue, i, P.sample_offset, ps, gd, g, index);
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0ull) break;
        if (alive && path_step<VOLUMES, TEXTURES, ENV, DISPERSION>(C, g, stack, ps, gd)) {
            PixelSums sums;
            sums.clear();
            sums.add(P, ps.result, gd, true);
            sums.store(P, index, accum, guide_albedo, guide_normal, comp);
            alive = false;
        }
    }
}

template <bool V, bool T, bool E, bool D, int WAVES>
static void launch_split_v(const LaunchMain& L, hipStream_t stream) {
    const uint32_t px = L.P.width * L.P.height;
    const SlotStore queue{L.split, px};
    uint32_t* counts = L.split + split_count_offset(px);
    launch_fill_u32(counts, 0u, 2, stream);
    with_tree(L.tree, [&](auto tree) {
        constexpr int TR = decltype(tree)::value;
        hipLaunchKernelGGL((k_split_head<V, T, E, D, TR, WAVES>), dim3((L.P.width + 15) / 16, (L.P.height + 15) / 16),
                           dim3(BLOCK_THREADS), 0, stream, L.S, L.P, L.accum, L.galb, L.gnorm, L.comp, L.spill,
                           L.slots, queue, counts, L.split_depth);
        hipLaunchKernelGGL((k_split_tail<V, T, E, D, TR, WAVES>), dim3(L.split_blocks), dim3(BLOCK_THREADS), 0, stream,
                           L.S, L.P, L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots, queue, counts);
    });
}

template <int WAVES>
static void launch_split_waves(const LaunchMain& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_split_v<false, false, false, false, WAVES>(L, stream); break;
    case 1: launch_split_v<false, false, false, true, WAVES>(L, stream); break;
    case 2: launch_split_v<false, false, true, false, WAVES>(L, stream); break;
    case 3: launch_split_v<false, false, true, true, WAVES>(L, stream); break;
    case 8: launch_split_v<true, false, false, false, WAVES>(L, stream); break;
    case 10: launch_split_v<true, false, true, false, WAVES>(L, stream); break;
    default: launch_split_v<true, true, true, true, WAVES>(L, stream); break;
    }
}

}  // namespace rm
