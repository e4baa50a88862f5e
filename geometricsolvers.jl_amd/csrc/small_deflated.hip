// libgsolve_hip: the lane-per-system kernels of a Deflated batch (src/nonlinear/deflation.jl:714-760):
// each family wrapped in FamDeflatedT, with the Static, Backtracking and DogLeg globalizations and
// `reuse = k` under Static and Backtracking, in a translation unit of its own.
#include "small_launch.hpp"

namespace gs {

namespace {

template <class Inner, int GLOB>
int launch_deflated_fam(bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg, hipStream_t st) {
    using Fam = FamDeflatedT<Inner>;
    if constexpr (GLOB != kDogLeg) {
        if (reuse) return launch_glob<Fam, GLOB, true>(a, mixed, refine, diag, wg, st);
    }
    return launch_glob<Fam, GLOB, false>(a, mixed, refine, diag, wg, st);
}

template <int GLOB>
int launch_deflated_glob(int family, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                         hipStream_t st) {
    switch (family) {
        case GS_FAM_COUPLED2:
            return launch_deflated_fam<FamCoupled2T<double>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_HETERO2:
            return launch_deflated_fam<FamHetero2T<double>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_SCALAR_SQUARE:
            return launch_deflated_fam<FamScalarSquareT<double>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_TANH_LINEAR3:
            return launch_deflated_fam<FamTanhLinear3T<double>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_HENON_HEILES_MIDPOINT:
            return launch_deflated_fam<FamHenonHeilesMidpointT<double>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a Deflated batch wraps one of the lane-per-system families (1-5)");
    }
}

}  // namespace

int launch_small_deflated(int family, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag,
                          int wg, hipStream_t st) {
    if (glob == kBacktracking) return launch_deflated_glob<kBacktracking>(family, reuse, a, mixed, refine, diag, wg, st);
    if (glob == kDogLeg) return launch_deflated_glob<kDogLeg>(family, false, a, mixed, refine, diag, wg, st);
    return launch_deflated_glob<kStatic>(family, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
