// libgsolve_hip: the lane-per-system kernels of a Reparametrized batch
// (src/nonlinear/reparametrized.jl:981-1031): a family wrapped in FamReparamT with a registered chart,
// the Static and Backtracking globalizations, in a translation unit of its own.
#include "small_launch.hpp"

namespace gs {

namespace {

template <int GLOB>
int launch_reparam_glob(int family, int chart, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                        hipStream_t st) {
    if (chart == GS_CHART_POLAR) {
        switch (family) {
            case GS_FAM_COUPLED2:
                return launch_glob<FamReparamT<FamCoupled2T<double>, ChartPolarT>, GLOB, false>(a, mixed, refine, diag,
                                                                                              wg, st);
            case GS_FAM_HETERO2:
                return launch_glob<FamReparamT<FamHetero2T<double>, ChartPolarT>, GLOB, false>(a, mixed, refine, diag,
                                                                                             wg, st);
            default: return set_error(GS_EFAMILY, "the polar chart maps 2 coordinates: families 1 and 2");
        }
    }
    switch (family) {
        case GS_FAM_COUPLED2:
            return launch_glob<FamReparamT<FamCoupled2T<double>, ChartLinearT<2>>, GLOB, false>(a, mixed, refine, diag,
                                                                                               wg, st);
        case GS_FAM_HETERO2:
            return launch_glob<FamReparamT<FamHetero2T<double>, ChartLinearT<2>>, GLOB, false>(a, mixed, refine, diag,
                                                                                              wg, st);
        case GS_FAM_SCALAR_SQUARE:
            return launch_glob<FamReparamT<FamScalarSquareT<double>, ChartLinearT<1>>, GLOB, false>(a, mixed, refine,
                                                                                                   diag, wg, st);
        case GS_FAM_TANH_LINEAR3:
            return launch_glob<FamReparamT<FamTanhLinear3T<double>, ChartLinearT<3>>, GLOB, false>(a, mixed, refine,
                                                                                                  diag, wg, st);
        case GS_FAM_HENON_HEILES_MIDPOINT:
            return launch_glob<FamReparamT<FamHenonHeilesMidpointT<double>, ChartLinearT<4>>, GLOB, false>(
                a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a Reparametrized batch wraps one of the lane-per-system families (1-5)");
    }
}

}  // namespace

int launch_small_reparam(int family, int chart, int glob, const KernelArgs& a, bool mixed, bool refine, bool diag,
                         int wg, hipStream_t st) {
    if (glob == kBacktracking) return launch_reparam_glob<kBacktracking>(family, chart, a, mixed, refine, diag, wg, st);
    return launch_reparam_glob<kStatic>(family, chart, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
