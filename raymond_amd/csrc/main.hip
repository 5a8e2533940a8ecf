// The path pass's launcher: picks the kernel for a frame's schedule, and
// k_main at four waves per SIMD (main.hpp).
#include "main.hpp"

namespace rm {

void launch_main(const LaunchMain& L, const Features& f, hipStream_t stream) {
    if (L.persistent_blocks > 0 && L.P.view == VIEW_COLOR) return launch_main_persistent(L, f, stream);
    if (L.staging && L.P.view == VIEW_COLOR) return launch_main_compact(L, f, stream);
    if (L.split && L.P.view == VIEW_COLOR && L.P.samples_per_frame == 1u) return launch_main_split(L, f, stream);
    if (L.waves == 5) return launch_main_waves5(L, f, stream);
    if (L.stack_depth == 16) return launch_main_stack16(L, f, stream);
    if (L.block_shape == BLOCK_ROWS) return launch_main_rows(L, f, stream);
    launch_main_pixel<4>(L, f, stream);
}

}  // namespace rm
