// libgsolve_hip: the lane-per-system kernels of a StationarityProblem batch
// (src/nonlinear/stationarity.jl:642-698): a registered action wrapped in FamStationarityT, with |E|
// = 2 or 6 unknowns, the Static, Backtracking and DogLeg globalizations and reuse = k, in a translation
// unit of its own.
#include "small_launch.hpp"

namespace gs {

namespace {

template <int GLOB, int NE>
int launch_stationarity_ne(int action, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                           hipStream_t st) {
    using Fam = FamStationarityT<ActHenonHeilesT<double>, NE>;
    switch (action) {
        case GS_ACT_HENON_HEILES:
            if constexpr (GLOB != kDogLeg) {
                if (reuse) return launch_glob<Fam, GLOB, true>(a, mixed, refine, diag, wg, st);
            }
            return launch_glob<Fam, GLOB, false>(a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "a StationarityProblem batch takes a registered action");
    }
}

template <int GLOB>
int launch_stationarity_glob(int action, int ne, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag,
                             int wg, hipStream_t st) {
    if (ne == 2) return launch_stationarity_ne<GLOB, 2>(action, reuse, a, mixed, refine, diag, wg, st);
    if (ne == 6) return launch_stationarity_ne<GLOB, 6>(action, reuse, a, mixed, refine, diag, wg, st);
    return set_error(GS_EFAMILY, "the device StationarityProblem batch takes |E| = 2 or 6 equations");
}

}  // namespace

int launch_small_stationarity(int action, int ne, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                              bool diag, int wg, hipStream_t st) {
    if (glob == kBacktracking)
        return launch_stationarity_glob<kBacktracking>(action, ne, reuse, a, mixed, refine, diag, wg, st);
    if (glob == kDogLeg) return launch_stationarity_glob<kDogLeg>(action, ne, false, a, mixed, refine, diag, wg, st);
    return launch_stationarity_glob<kStatic>(action, ne, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
