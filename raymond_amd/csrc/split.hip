// The split schedule's launcher and its kernels at four waves per SIMD
// (split.hpp); split5.hip holds the five-wave instances.
#include "split.hpp"

namespace rm {

void launch_main_split(const LaunchMain& L, const Features& f, hipStream_t stream) {
    if (L.waves == 5) return launch_main_split5(L, f, stream);
    launch_split_waves<4>(L, f, stream);
}

}  // namespace rm
