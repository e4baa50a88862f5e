// The launch of the mid-size solve (one wave per system): the arguments gs_solve_batch_f64 hands to
// the wave-per-system kernels of mid_solver.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gsolve.h"
#include "small_solver.hpp"

namespace gs {

// The kernels' view of a method and its options (defined with gs_solve_batch_f64).
DevParams dev_params(const gs_method& method, const gs_options_f64& options);

// The gs_set_kernel_timing bracket around one launch (defined with gs_solve_batch_f64).
bool timing_begin(hipStream_t st);
bool timing_mark(hipStream_t st, bool timed);
void timing_end(bool timed);

struct MidLaunch {
    DevParams q;
    const char* P;        // parameter records
    int64_t p_stride;     // bytes between records (0: one shared record)
    double shared[8];     // the family's shared scalars
    int64_t nsys;
    double* X;
    gs_status_f64* S;
    uint8_t* codes;
    const int64_t* order;
    bool mixed, refine;
    int32_t glob;         // kStatic, kBacktracking or kDogLeg
    bool reuse;           // reuse = k (the period in q.reuse)
    uint8_t* redo;        // the hand-over flags between the two kernels, nsys bytes; null: the
                          // library's own (launch_fpu_gl3 sets them)
};

// The FPU-β Gauss–Legendre s = 3 stage system (GS_FAM_FPU_GL3, n = 48) with the Static, the
// Backtracking (with or without reuse = k) or the DogLeg globalization; returns a gs_error code (the launch only;
// the caller synchronises).
int launch_fpu_gl3(const MidLaunch& L, hipStream_t st);

// The reuse = k forms of launch_fpu_gl3 (L.redo set), in their own translation unit.
int launch_fpu_gl3_reuse(const MidLaunch& L, hipStream_t st);

}  // namespace gs
