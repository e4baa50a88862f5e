// libgsolve_hip: the lane-per-system kernels of Float32 batches (gs_solve_batch_f32), every
// globalization and reuse = k, in a translation unit of their own.
#include "small_launch.hpp"

namespace gs {

int launch_small_f32(int family, int glob, bool reuse, const KernelArgsT<float>& a, bool refine, bool diag, int wg,
                     hipStream_t st) {
    if (glob == GS_DOGLEG) return launch_small_family<kDogLeg, false>(family, a, false, refine, diag, wg, st);
    if (glob == GS_STATIC)
        return reuse ? launch_small_family<kStatic, true>(family, a, false, refine, diag, wg, st)
                     : launch_small_family<kStatic, false>(family, a, false, refine, diag, wg, st);
    return reuse ? launch_small_family<kBacktracking, true>(family, a, false, refine, diag, wg, st)
                 : launch_small_family<kBacktracking, false>(family, a, false, refine, diag, wg, st);
}

}  // namespace gs
