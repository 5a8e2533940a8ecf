// k_main at four waves per SIMD with the tuner's other stack depth and block
// shape (main.hpp), in a file of its own so that it compiles beside the
// default instances.
#include "main.hpp"

namespace rm {

void launch_main_stack16(const LaunchMain& L, const Features& f, hipStream_t stream) {
    launch_main_pixel<4, 16, BLOCK_SQUARES>(L, f, stream);
}

void launch_main_rows(const LaunchMain& L, const Features& f, hipStream_t stream) {
    launch_main_pixel<4, MAIN_LDS_DEPTH, BLOCK_ROWS>(L, f, stream);
}

}  // namespace rm
