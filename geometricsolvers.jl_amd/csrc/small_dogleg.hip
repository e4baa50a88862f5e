// libgsolve_hip: the lane-per-system kernels of Newton(; globalization = DogLeg()) (the trial loop of
// dogleg.jl:116-206 on the factorised Jacobian, newton.jl:1725-1780), in a translation unit of
// their own.
#include "small_launch.hpp"

namespace gs {

int launch_small_dogleg(int family, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg, hipStream_t st) {
    return launch_small_family<kDogLeg, false>(family, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
