// The launch of the dense mid-size solve (a segment of 16, 32 or 64 lanes per system, n = 16, 24, 32,
// 48 or 64): the arguments gs_solve_batch_f64 hands to the kernels of dense_solver.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gsolve.h"
#include "small_solver.hpp"

namespace gs {

struct DenseLaunch {
    DevParams q;
    const double* A;      // the family's matrix (n × n, column-major), device
    const char* P;        // parameter records (b, n doubles each)
    int64_t p_stride;     // bytes between records (0: one shared record)
    int64_t nsys;
    double* X;
    gs_status_f64* S;
    uint8_t* codes;
    const int64_t* order;
};

// The systems of one launch per workgroup at size n with its current lanes per system.
int dense_systems_per_block(int n);

// The lanes per system (segment width: 16, 32 or 64, at least n; 0 = the default, the smallest that
// holds n rows) of the launches at size n from now on; GS_EINVAL for another size or width.
int dense_set_shape(int n, int lanes);

// The lanes per system the next launch at size n uses.
int lanes_of(int n);

// GS_FAM_DENSE_SIN (n = 16, 24, 32, 48 or 64) under Newton with the Static or Backtracking
// globalization, with or without reuse = k, LU in FP64 or FP32 (mixed, with or without Refinement);
// returns a gs_error code (the launch only; the caller synchronises).
int launch_dense(int n, int glob, bool mixed, bool refine, bool diag, bool reuse, const DenseLaunch& L,
                 hipStream_t st);

}  // namespace gs
