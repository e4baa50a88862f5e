// The dense mid-size solve at n = 24 under DogLeg (dense_kernel.hpp), lanes per system 32 or 64.
#include "dense_kernel.hpp"

namespace gs {

int launch_dense_dogleg_n24(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st) {
    if (np == 32) return launch_np_dogleg<24, 32>(mixed, refine, diag, L, st);
    if (np == 64) return launch_np_dogleg<24, 64>(mixed, refine, diag, L, st);
    return set_error(GS_EINVAL, "no dense launch with these lanes per system at n = 24");
}

}  // namespace gs
