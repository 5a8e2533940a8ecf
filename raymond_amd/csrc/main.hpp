// k_main, the default path-tracing pass (one thread per pixel), and its
// launch for a frame's features, at a given number of waves per SIMD.
// main.hip instantiates it at four, main5.hip at five, so that the two
// compile in parallel.
#pragma once
#include "path.hpp"

namespace rm {

// src/shaders.mjs:1838-1895, one thread per pixel: the frame's samples one
// after another, each path run to its end.
//
// WAVES is the occupancy the compiler is held to. Left to itself it keeps
// every value of a path in registers (153 VGPRs, three waves). At four waves
// (128 VGPRs) it spills about 100 bytes a lane of rarely touched state (the
// pixel's sums, the guides), and the extra wave hides enough traversal
// latency to make the pass about 11% faster on the 1080p sphere scene. Five
// (96 VGPRs) spills more and is faster again alone: 12% on a 2M-triangle
// scene and on glass caustics, whose frames the path pass dominates. On the
// sphere scene, whose light tracing and gather take half as long as the
// path pass, the fifth wave crowds them out and the frame is 1% slower.
// RM_OPTION_WAVES chooses; rm_tune tries both.
//
// LDS_DEPTH (RM_OPTION_STACK_DEPTH) is how many traversal stack entries stay
// in LDS, the rest spilling to global memory: 8 (16 KiB a block) or 16
// (32 KiB). SHAPE (RM_OPTION_BLOCK_SHAPE) maps the block's threads to
// pixels: 16×16 squares or 64×4 rows. The tuner tries both.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE, int WAVES, int LDS_DEPTH = MAIN_LDS_DEPTH,
          int SHAPE = BLOCK_SQUARES>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_main(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots) {
    __shared__ uint2 lds[LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    shaped_pixel<SHAPE>(x, y);
    main_pixel<VOLUMES, TEXTURES, ENV, DISPERSION, TREE, LDS_DEPTH>(S, P, x, y, accum, guide_albedo, guide_normal, comp,
                                                                    spill, slots, lds);
}

template <bool V, bool T, bool E, bool D, int WAVES, int LDS_DEPTH, int SHAPE>
static void launch_main_v(const LaunchMain& L, hipStream_t stream) {
    const dim3 grid = shaped_grid<SHAPE>(L.P.width, L.P.height);
    with_tree(L.tree, [&](auto tree) {
        hipLaunchKernelGGL((k_main<V, T, E, D, decltype(tree)::value, WAVES, LDS_DEPTH, SHAPE>), grid,
                           dim3(BLOCK_THREADS), 0, stream, L.S, L.P, L.accum, L.galb, L.gnorm, L.comp, L.spill,
                           L.slots);
    });
}

template <int WAVES, int LDS_DEPTH = MAIN_LDS_DEPTH, int SHAPE = BLOCK_SQUARES>
static void launch_main_pixel(const LaunchMain& L, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_main_v<false, false, false, false, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    case 1: launch_main_v<false, false, false, true, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    case 2: launch_main_v<false, false, true, false, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    case 3: launch_main_v<false, false, true, true, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    case 8: launch_main_v<true, false, false, false, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    case 10: launch_main_v<true, false, true, false, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    default: launch_main_v<true, true, true, true, WAVES, LDS_DEPTH, SHAPE>(L, stream); break;
    }
}

}  // namespace rm
