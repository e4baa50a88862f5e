// The dense mid-size solve of a Float32 batch at n = 48 (dense_kernel.hpp), lanes per system 64,
// with and without reuse = k.
#include "dense_kernel.hpp"

namespace gs {

int launch_dense_f32_n48(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st) {
    if (np == 64) return launch_np_f32<48, 64>(glob, refine, diag, reuse, L, st);
    return set_error(GS_EINVAL, "no dense launch with these lanes per system at n = 48");
}

}  // namespace gs
