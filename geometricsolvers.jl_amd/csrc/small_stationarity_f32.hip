// libgsolve_hip: the lane-per-system kernels of a Float32 StationarityProblem batch, every form of
// the Float64 batch but the mixed one, in a translation unit of its own.
#include "wrapper_launch.hpp"

namespace gs {

int launch_small_stationarity_f32(int action, int ne, int glob, bool reuse, const KernelArgsT<float>& a, bool refine,
                                  bool diag, int wg, hipStream_t st) {
    return launch_stationarity_t<float>(action, ne, glob, reuse, a, false, refine, diag, wg, st);
}

}  // namespace gs
