// k_main_batch: several jobs' path passes in one launch (rm_frame_batch).
#include "path.hpp"

namespace rm {

// Several jobs' path passes in one launch (rm_frame_batch): the grid's 16×16
// tiles are the jobs' tiles one job after another, and each block renders
// its tile of its job exactly as k_main would. The job table is the kernel
// argument, so a block's job (the same for all its lanes) is read by scalar
// loads from the argument segment.
template <bool VOLUMES, bool TEXTURES, bool ENV, bool DISPERSION, int TREE>
__global__ __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_main_batch(BatchTable T) {
    __shared__ uint2 lds[MAIN_LDS_DEPTH * BLOCK_THREADS];
    uint32_t j = 0;
    while (j + 1u < T.n && blockIdx.x >= T.begin[j + 1u]) ++j;
    const BatchJob& J = T.job[j];
    const uint32_t tile = blockIdx.x - T.begin[j];
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x = (tile % J.tiles_x) * 16u + (w & 1u) * 8u + (l & 7u);
    const uint32_t y = (tile / J.tiles_x) * 16u + (w >> 1) * 8u + (l >> 3);
    main_pixel<VOLUMES, TEXTURES, ENV, DISPERSION, TREE>(J.S, J.P, x, y, J.accum, J.galb, J.gnorm, J.comp, J.spill,
                                                         J.slots, lds);
}

// ------------------------------------------------------------------ launchers
template <bool V, bool T, bool E, bool D>
static void launch_main_batch_v(const BatchTable& B, int tree, hipStream_t stream) {
    with_tree(tree, [&](auto t) {
        hipLaunchKernelGGL((k_main_batch<V, T, E, D, decltype(t)::value>), dim3(B.begin[B.n]), dim3(BLOCK_THREADS), 0,
                           stream, B);
    });
}

void launch_main_batch(const BatchTable& B, const Features& f, hipStream_t stream) {
    int key = (f.volumes ? 8 : 0) | (f.textures ? 4 : 0) | (f.environment ? 2 : 0) | (f.dispersion ? 1 : 0);
    switch (key) {
    case 0: launch_main_batch_v<false, false, false, false>(B, f.tree, stream); break;
    case 1: launch_main_batch_v<false, false, false, true>(B, f.tree, stream); break;
    case 2: launch_main_batch_v<false, false, true, false>(B, f.tree, stream); break;
    case 3: launch_main_batch_v<false, false, true, true>(B, f.tree, stream); break;
    case 8: launch_main_batch_v<true, false, false, false>(B, f.tree, stream); break;
    case 10: launch_main_batch_v<true, false, true, false>(B, f.tree, stream); break;
    default: launch_main_batch_v<true, true, true, true>(B, f.tree, stream); break;
    }
}

}  // namespace rm
