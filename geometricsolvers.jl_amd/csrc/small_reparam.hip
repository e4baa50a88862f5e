// libgsolve_hip: the lane-per-system kernels of a Reparametrized batch
// (src/nonlinear/reparametrized.jl:981-1031): a family wrapped in FamReparamT with a registered chart,
// the Static, Backtracking and DogLeg globalizations and `reuse = k` under Static and Backtracking,
// in a translation unit of its own.
#include "small_launch.hpp"

namespace gs {

namespace {

template <class Fam, int GLOB>
int launch_reparam_fam(bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg, hipStream_t st) {
    if constexpr (GLOB != kDogLeg) {
        if (reuse) return launch_glob<Fam, GLOB, true>(a, mixed, refine, diag, wg, st);
    }
    return launch_glob<Fam, GLOB, false>(a, mixed, refine, diag, wg, st);
}

template <int GLOB>
int launch_reparam_glob(int family, int chart, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag,
                        int wg, hipStream_t st) {
    if (chart == GS_CHART_POLAR) {
        switch (family) {
            case GS_FAM_COUPLED2:
                return launch_reparam_fam<FamReparamT<FamCoupled2T<double>, ChartPolarT>, GLOB>(reuse, a, mixed, refine,
                                                                                               diag, wg, st);
            case GS_FAM_HETERO2:
                return launch_reparam_fam<FamReparamT<FamHetero2T<double>, ChartPolarT>, GLOB>(reuse, a, mixed, refine,
                                                                                              diag, wg, st);
            default: return set_error(GS_EFAMILY, "the polar chart maps 2 coordinates: families 1 and 2");
        }
    }
    switch (family) {
        case GS_FAM_COUPLED2:
            return launch_reparam_fam<FamReparamT<FamCoupled2T<double>, ChartLinearT<2>>, GLOB>(reuse, a, mixed, refine,
                                                                                                diag, wg, st);
        case GS_FAM_HETERO2:
            return launch_reparam_fam<FamReparamT<FamHetero2T<double>, ChartLinearT<2>>, GLOB>(reuse, a, mixed, refine,
                                                                                               diag, wg, st);
        case GS_FAM_SCALAR_SQUARE:
            return launch_reparam_fam<FamReparamT<FamScalarSquareT<double>, ChartLinearT<1>>, GLOB>(
                reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_TANH_LINEAR3:
            return launch_reparam_fam<FamReparamT<FamTanhLinear3T<double>, ChartLinearT<3>>, GLOB>(
                reuse, a, mixed, refine, diag, wg, st);
        case GS_FAM_HENON_HEILES_MIDPOINT:
            return launch_reparam_fam<FamReparamT<FamHenonHeilesMidpointT<double>, ChartLinearT<4>>, GLOB>(
                reuse, a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a Reparametrized batch wraps one of the lane-per-system families (1-5)");
    }
}

}  // namespace

int launch_small_reparam(int family, int chart, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                         bool diag, int wg, hipStream_t st) {
    if (glob == kBacktracking)
        return launch_reparam_glob<kBacktracking>(family, chart, reuse, a, mixed, refine, diag, wg, st);
    if (glob == kDogLeg) return launch_reparam_glob<kDogLeg>(family, chart, false, a, mixed, refine, diag, wg, st);
    return launch_reparam_glob<kStatic>(family, chart, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
