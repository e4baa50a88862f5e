// libgsolve_hip: the lane-per-system kernels of a Float32 Reparametrized batch with the linear chart,
// every form of the Float64 batch but the mixed one, in a translation unit of its own.
#include "wrapper_launch.hpp"

namespace gs {

int launch_small_reparam_f32(int family, int chart, int glob, bool reuse, const KernelArgsT<float>& a, bool refine,
                             bool diag, int wg, hipStream_t st) {
    return launch_reparam_t<float>(family, chart, glob, reuse, a, false, refine, diag, wg, st);
}

}  // namespace gs
