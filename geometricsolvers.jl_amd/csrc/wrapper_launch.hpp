// The launchers of the StationarityProblem and Reparametrized batches, templated on the working
// precision T: small_stationarity.hip and small_reparam.hip instantiate them in Float64,
// small_stationarity_f32.hip and small_reparam_f32.hip in Float32 (whose factor is in the working
// precision already, so `mixed` is ignored there).
#pragma once

#include "small_launch.hpp"

namespace gs {

// One family's kernels under one globalization, with or without reuse = k (DogLeg takes none).
template <class Fam, int GLOB>
int launch_wrapped(bool reuse, const KernelArgsT<typename Fam::T>& a, bool mixed, bool refine, bool diag, int wg,
                   hipStream_t st) {
    if constexpr (GLOB != kDogLeg) {
        if (reuse) return launch_glob<Fam, GLOB, true>(a, mixed, refine, diag, wg, st);
    }
    return launch_glob<Fam, GLOB, false>(a, mixed, refine, diag, wg, st);
}

// A registered action wrapped in FamStationarityT with |E| = NE unknowns.
template <class T, int GLOB, int NE>
int launch_stationarity_ne(int action, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine, bool diag,
                           int wg, hipStream_t st) {
    switch (action) {
        case GS_ACT_HENON_HEILES:
            return launch_wrapped<FamStationarityT<ActHenonHeilesT<T>, NE>, GLOB>(reuse, a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a StationarityProblem batch takes a registered action");
    }
}

template <class T, int GLOB>
int launch_stationarity_glob(int action, int ne, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine,
                             bool diag, int wg, hipStream_t st) {
    if (ne == 2) return launch_stationarity_ne<T, GLOB, 2>(action, reuse, a, mixed, refine, diag, wg, st);
    if (ne == 6) return launch_stationarity_ne<T, GLOB, 6>(action, reuse, a, mixed, refine, diag, wg, st);
    return set_error(GS_EFAMILY, "the device StationarityProblem batch takes |E| = 2 or 6 equations");
}

template <class T>
int launch_stationarity_t(int action, int ne, int glob, bool reuse, const KernelArgsT<T>& a, bool mixed,
                          bool refine, bool diag, int wg, hipStream_t st) {
    if (glob == kBacktracking)
        return launch_stationarity_glob<T, kBacktracking>(action, ne, reuse, a, mixed, refine, diag, wg, st);
    if (glob == kDogLeg) return launch_stationarity_glob<T, kDogLeg>(action, ne, false, a, mixed, refine, diag, wg, st);
    return launch_stationarity_glob<T, kStatic>(action, ne, reuse, a, mixed, refine, diag, wg, st);
}

// A lane-per-system family wrapped in FamReparamT: the linear chart for families 1–5, the polar
// chart (Float64 only) for the n = 2 families.
template <class T, int GLOB>
int launch_reparam_glob(int family, int chart, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine,
                        bool diag, int wg, hipStream_t st) {
    if (chart == GS_CHART_POLAR) {
        if constexpr (sizeof(T) == 8) {
            switch (family) {
                case GS_FAM_COUPLED2:
                    return launch_wrapped<FamReparamT<FamCoupled2T<T>, ChartPolarT>, GLOB>(reuse, a, mixed, refine,
                                                                                           diag, wg, st);
                case GS_FAM_HETERO2:
                    return launch_wrapped<FamReparamT<FamHetero2T<T>, ChartPolarT>, GLOB>(reuse, a, mixed, refine,
                                                                                          diag, wg, st);
                default: return set_error(GS_EFAMILY, "the polar chart maps 2 coordinates: families 1 and 2");
            }
        }
        return set_error(GS_EFAMILY, "the Float32 Reparametrized batch takes the linear chart");
    }
    switch (family) {
        case GS_FAM_COUPLED2:
            return launch_wrapped<FamReparamT<FamCoupled2T<T>, ChartLinearT<2, T>>, GLOB>(reuse, a, mixed, refine, diag,
                                                                                         wg, st);
        case GS_FAM_HETERO2:
            return launch_wrapped<FamReparamT<FamHetero2T<T>, ChartLinearT<2, T>>, GLOB>(reuse, a, mixed, refine, diag,
                                                                                        wg, st);
        case GS_FAM_SCALAR_SQUARE:
            return launch_wrapped<FamReparamT<FamScalarSquareT<T>, ChartLinearT<1, T>>, GLOB>(reuse, a, mixed, refine,
                                                                                             diag, wg, st);
        case GS_FAM_TANH_LINEAR3:
            return launch_wrapped<FamReparamT<FamTanhLinear3T<T>, ChartLinearT<3, T>>, GLOB>(reuse, a, mixed, refine,
                                                                                            diag, wg, st);
        case GS_FAM_HENON_HEILES_MIDPOINT:
            return launch_wrapped<FamReparamT<FamHenonHeilesMidpointT<T>, ChartLinearT<4, T>>, GLOB>(
                reuse, a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a Reparametrized batch wraps one of the lane-per-system families (1-5)");
    }
}

template <class T>
int launch_reparam_t(int family, int chart, int glob, bool reuse, const KernelArgsT<T>& a, bool mixed, bool refine,
                     bool diag, int wg, hipStream_t st) {
    if (glob == kBacktracking)
        return launch_reparam_glob<T, kBacktracking>(family, chart, reuse, a, mixed, refine, diag, wg, st);
    if (glob == kDogLeg) return launch_reparam_glob<T, kDogLeg>(family, chart, false, a, mixed, refine, diag, wg, st);
    return launch_reparam_glob<T, kStatic>(family, chart, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
