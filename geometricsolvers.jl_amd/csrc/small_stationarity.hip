// libgsolve_hip: the lane-per-system kernels of a StationarityProblem batch
// (src/nonlinear/stationarity.jl:642-698) in Float64: a registered action wrapped in FamStationarityT,
// with |E| = 2 or 6 unknowns, the Static, Backtracking and DogLeg globalizations and reuse = k, in a
// translation unit of its own.
#include "wrapper_launch.hpp"

namespace gs {

int launch_small_stationarity(int action, int ne, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                              bool diag, int wg, hipStream_t st) {
    return launch_stationarity_t<double>(action, ne, glob, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
