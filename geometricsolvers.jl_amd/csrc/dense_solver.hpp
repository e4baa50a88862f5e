// The launch of the dense mid-size solve (a segment of 16, 32 or 64 lanes per system, n = 16, 24, 32,
// 48 or 64): the arguments gs_solve_batch_f64 hands to the kernels of dense_solver.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/gsolve.h"
#include "small_solver.hpp"

namespace gs {

// In the working precision T (double, or float for a Float32 batch).
template <class T> struct DenseLaunchT {
    DevParamsT<T> q;
    const T* A;           // the family's matrix (n × n, column-major), device
    const char* P;        // parameter records (b, n entries of T each)
    int64_t p_stride;     // bytes between records (0: one shared record)
    int64_t nsys;
    T* X;
    typename std::conditional<sizeof(T) == 8, gs_status_f64, gs_status_f32>::type* S;
    uint8_t* codes;
    const int64_t* order;
};
using DenseLaunch = DenseLaunchT<double>;

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

// The same in Float32 (a Float32 batch: the factor is in the working precision, so no mixed form).
int launch_dense_f32(int n, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                     hipStream_t st);

}  // namespace gs
