// libgsolve_hip, mid-size systems: the wave-per-system kernels of n = 48 with reuse = k (the held
// factor's rows kept in VGPRs across steps), in a translation unit of their own.
#include "mid_kernel.hpp"

namespace gs {

int launch_fpu_gl3_reuse(const MidLaunch& L, hipStream_t st) { return launch_forms<true>(L, st); }

}  // namespace gs
