// libgsolve_hip: the lane-per-system kernels of a Reparametrized batch
// (src/nonlinear/reparametrized.jl:981-1031) in Float64: a family wrapped in FamReparamT with a
// registered chart, the Static, Backtracking and DogLeg globalizations and `reuse = k` under Static
// and Backtracking, in a translation unit of its own.
#include "wrapper_launch.hpp"

namespace gs {

int launch_small_reparam(int family, int chart, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                         bool diag, int wg, hipStream_t st) {
    return launch_reparam_t<double>(family, chart, glob, reuse, a, mixed, refine, diag, wg, st);
}

}  // namespace gs
