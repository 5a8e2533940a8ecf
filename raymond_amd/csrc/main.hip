// k_main: the default path-tracing pass, one thread per pixel
// (src/shaders.mjs:1838-1895), and the launcher that picks the pass for a
// frame's schedule.
#include "path.hpp"

namespace rm {

// ------------------------------------------------------------------ main
// src/shaders.mjs:1838-1895, one thread per pixel: the frame's samples one
// after another, each path run to its end.
//
// Four waves per SIMD: left to itself the compiler keeps every value of a
// path in registers (153 VGPRs, three waves); capping it at 128 spills about
// 112 bytes a lane of rarely touched state (the pixel's sums, the guides) and
// the extra wave hides enough traversal latency to make the pass about 11%
// faster on the 1080p sphere scene.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main(
    DevScene S, Params P, float4* __restrict__ accum, float4* __restrict__ guide_albedo,
    float4* __restrict__ guide_normal, float4* __restrict__ comp, uint32_t* spill, uint32_t slots) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t x, y;
    tile_pixel(x, y);
    main_pixel<VOLUMES, TEXTURES, ENV, DISPERSION, TREE>(S, P, x, y, accum, guide_albedo, guide_normal, comp, spill,
                                                         slots, lds);
}

// ------------------------------------------------------------------ launchers
template <bool V, bool T, bool E, bool D>
static void launch_main_v(const LaunchMain& L, hipStream_t stream) {
    dim3 grid((L.P.width + 15) / 16, (L.P.height + 15) / 16);
    with_tree(L.tree, [&](auto tree) {
        hipLaunchKernelGGL((k_main<V, T, E, D, decltype(tree)::value>), grid, dim3(BLOCK_THREADS), 0, stream, L.S, L.P,
                           L.accum, L.galb, L.gnorm, L.comp, L.spill, L.slots);
    });
}

void launch_main(const LaunchMain& L, const Features& f, hipStream_t stream) {
    if (L.persistent_blocks > 0 && L.P.view == VIEW_COLOR) return launch_main_persistent(L, f, stream);
    if (L.staging && L.P.view == VIEW_COLOR) return launch_main_compact(L, f, stream);
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_main_v<false, false, false, false>(L, stream); break;
    case 1: launch_main_v<false, false, false, true>(L, stream); break;
    case 2: launch_main_v<false, false, true, false>(L, stream); break;
    case 3: launch_main_v<false, false, true, true>(L, stream); break;
    case 8: launch_main_v<true, false, false, false>(L, stream); break;
    case 10: launch_main_v<true, false, true, false>(L, stream); break;
    default: launch_main_v<true, true, true, true>(L, stream); break;
    }
}

}  // namespace rm
