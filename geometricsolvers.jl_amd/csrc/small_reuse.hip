// libgsolve_hip: the lane-per-system kernels of Newton(; reuse = k) (newton.jl:1781-1874) with the
// Static and Backtracking globalizations, in a translation unit of their own.
#include "small_launch.hpp"

namespace gs {

int launch_small_reuse(int family, int glob, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                       hipStream_t st) {
    if (glob == GS_STATIC) return launch_small_family<kStatic, true>(family, a, mixed, refine, diag, wg, st);
    return launch_small_family<kBacktracking, true>(family, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
