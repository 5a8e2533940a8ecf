// k_main at five waves per SIMD (main.hpp), in a file of its own so that it
// compiles beside the four-wave instances.
#include "main.hpp"

namespace rm {

void launch_main_waves5(const LaunchMain& L, const Features& f, hipStream_t stream) {
    launch_main_pixel<5>(L, f, stream);
}

}  // namespace rm
