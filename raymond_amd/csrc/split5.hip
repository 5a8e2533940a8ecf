// The split schedule's kernels at five waves per SIMD (split.hpp), in a file
// of their own so that they compile beside the four-wave instances.
#include "split.hpp"

namespace rm {

void launch_main_split5(const LaunchMain& L, const Features& f, hipStream_t stream) {
    launch_split_waves<5>(L, f, stream);
}

}  // namespace rm
