// The dense mid-size solve of a Float32 batch at n = 12 (dense_kernel.hpp), lanes per system 16 or 32,
// with and without reuse = k.
#include "dense_kernel.hpp"

namespace gs {

int launch_dense_f32_n12(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st) {
    if (np == 16) return launch_np_f32<12, 16>(glob, refine, diag, reuse, L, st);
    if (np == 32) return launch_np_f32<12, 32>(glob, refine, diag, reuse, L, st);
    return set_error(GS_EINVAL, "no dense launch with these lanes per system at n = 12");
}

}  // namespace gs
