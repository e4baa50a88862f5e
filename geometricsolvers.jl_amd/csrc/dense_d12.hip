// The dense mid-size solve at n = 12 under DogLeg (dense_kernel.hpp), lanes per system 16 or 32.
#include "dense_kernel.hpp"

namespace gs {

int launch_dense_dogleg_n12(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st) {
    if (np == 16) return launch_np_dogleg<12, 16>(mixed, refine, diag, L, st);
    if (np == 32) return launch_np_dogleg<12, 32>(mixed, refine, diag, L, st);
    return set_error(GS_EINVAL, "no dense launch with these lanes per system at n = 12");
}

}  // namespace gs
