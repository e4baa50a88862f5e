// The dense mid-size solve at n = 64 under DogLeg (dense_kernel.hpp), lanes per system 64.
#include "dense_kernel.hpp"

namespace gs {

int launch_dense_dogleg_n64(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st) {
    if (np == 64) return launch_np_dogleg<64, 64>(mixed, refine, diag, L, st);
    return set_error(GS_EINVAL, "no dense launch with these lanes per system at n = 64");
}

}  // namespace gs
