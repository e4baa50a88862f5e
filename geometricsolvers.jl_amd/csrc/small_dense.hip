// libgsolve_hip: the dense family (GS_FAM_DENSE_SIN) at n = 2, 4, 6 and 8 (and 12 in Float32) on the
// lane-per-system kernels (FamDenseSinSmallT), in FP64 and in Float32, every globalization, with or
// without reuse = k (not with DogLeg), in a translation unit of its own.
#include "small_launch.hpp"

namespace gs {

namespace {

template <class T, int N>
int launch_dense_small_n(int glob, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine, bool diag, int wg,
                         hipStream_t st) {
    using Fam = FamDenseSinSmallT<T, N>;
    if (glob == kDogLeg) return launch_glob<Fam, kDogLeg, false>(a, mixed, refine, diag, wg, st);
    if (glob == kBacktracking)
        return reuse ? launch_glob<Fam, kBacktracking, true>(a, mixed, refine, diag, wg, st)
                     : launch_glob<Fam, kBacktracking, false>(a, mixed, refine, diag, wg, st);
    return reuse ? launch_glob<Fam, kStatic, true>(a, mixed, refine, diag, wg, st)
                 : launch_glob<Fam, kStatic, false>(a, mixed, refine, diag, wg, st);
}

template <class T>
int launch_dense_small_t(int n, int glob, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine, bool diag,
                         int wg, hipStream_t st) {
    switch (n) {
        case 2: return launch_dense_small_n<T, 2>(glob, reuse, a, mixed, refine, diag, wg, st);
        case 4: return launch_dense_small_n<T, 4>(glob, reuse, a, mixed, refine, diag, wg, st);
        case 6: return launch_dense_small_n<T, 6>(glob, reuse, a, mixed, refine, diag, wg, st);
        case 8: return launch_dense_small_n<T, 8>(glob, reuse, a, mixed, refine, diag, wg, st);
        case 12:
            // Float32 only: there the rows fit a lane and run faster than on the segment kernel; an
            // FP64 batch at n = 12 takes the segment kernel
            if constexpr (sizeof(T) == 4) return launch_dense_small_n<T, 12>(glob, reuse, a, mixed, refine, diag, wg, st);
            break;
    }
    return set_error(GS_EINVAL, "the lane-per-system dense family takes n = 2, 4, 6, 8 (and 12 in Float32)");
}

}  // namespace

int launch_small_dense(int n, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                       hipStream_t st) {
    return launch_dense_small_t<double>(n, glob, reuse, a, mixed, refine, diag, wg, st);
}

int launch_small_dense_f32(int n, int glob, bool reuse, const KernelArgsT<float>& a, bool refine, bool diag, int wg,
                           hipStream_t st) {
    return launch_dense_small_t<float>(n, glob, reuse, a, false, refine, diag, wg, st);
}

}  // namespace gs
