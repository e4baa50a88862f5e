// The registered problem families as device functions: residual F(x, p) and analytic Jacobian
// J(x, p) of one system, with its parameter record loaded into registers once per system. The Julia
// extension (ext/GeometricSolversHIPExt.jl) defines the same F and J as the `NonlinearProblem` of
// each family, and these follow the operation order of those Julia expressions (no contraction;
// `muladd` only where StaticArrays' matrix-vector product uses it).
//
// A family provides: N (unknowns), `load(args, idx)`, `F(x, r)` and `J(x, Jm)` with Jm[i][j] the
// entry (i, j).
#pragma once

#include "device_common.hpp"

namespace gs {

// Launch arguments shared by every family.
struct FamArgs {
    const char* P;       // parameter records (device)
    int64_t p_stride;    // bytes between records; 0 = one shared record
    double shared[8];    // scalar parameters shared by every system
};

GS_DEV const char* record(const FamArgs& a, int64_t idx) { return a.P + (a.p_stride ? idx * a.p_stride : 0); }

// The working precision of a family's records and arithmetic: T = double for gs_solve_batch_f64,
// float for gs_solve_batch_f32; constants are T(c), as the reference's generic residuals promote
// their literals to the element type.
GS_DEV double ttanh(double a) { return tanh(a); }
GS_DEV float ttanh(float a) { return tanhf(a); }

// _r2_coupled_res / _r2_coupled_jac (test/nonlinear/batched.jl:92-103).
template <class R> struct FamCoupled2T {
    using T = R;
    static constexpr int N = 2;
    T p0, p1;
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const T* p = reinterpret_cast<const T*>(record(a, idx));
        p0 = p[0];
        p1 = p[1];
    }
    GS_DEV void F(const T (&x)[2], T (&r)[2]) const {
        r[0] = (x[0] + (x[1] * x[1]) / T(8)) - p0;
        r[1] = (x[1] + (x[0] * x[0]) / T(8)) - p1;
    }
    GS_DEV void J(const T (&x)[2], T (&Jm)[2][2]) const {
        Jm[0][0] = T(1);
        Jm[0][1] = x[1] / T(4);
        Jm[1][0] = x[0] / T(4);
        Jm[1][1] = T(1);
    }
};

// _r2_hetero_res / _r2_hetero_jac (test/nonlinear/batched.jl:54-90): one fixture per ReturnCode.
// The record is an Int64 code in either precision.
template <class R> struct FamHetero2T {
    using T = R;
    static constexpr int N = 2;
    int64_t code;
    GS_DEV void load(const FamArgs& a, int64_t idx) { code = *reinterpret_cast<const int64_t*>(record(a, idx)); }
    GS_DEV void F(const T (&x)[2], T (&r)[2]) const {
        if (code == 1) {
            r[0] = x[0] - T(2);
            r[1] = x[1] - T(3);
        } else if (code == 2 || code == 4) {
            r[0] = T(1);
            r[1] = T(1);
        } else if (code == 3) {
            r[0] = x[0] - T(100);
            r[1] = x[1] - T(100);
        } else if (code == 5) {
            r[0] = tnan<T>();
            r[1] = tnan<T>();
        } else {
            r[0] = x[0] * x[0] + T(1);
            r[1] = x[1] * x[1] + T(1);
        }
    }
    GS_DEV void J(const T (&)[2], T (&Jm)[2][2]) const {
        T d = code == 1 ? T(1) : code == 2 ? T(1e16) : code == 3 ? T(2) : code == 4 ? T(0) : code == 5 ? T(1) : T(-1);
        Jm[0][0] = d;
        Jm[0][1] = T(0);
        Jm[1][0] = T(0);
        Jm[1][1] = d;
    }
};

// _r2_scalar_res / _r2_scalar_jac (test/nonlinear/batched.jl:107-108): x² − p, 2x.
template <class R> struct FamScalarSquareT {
    using T = R;
    static constexpr int N = 1;
    T p;
    GS_DEV void load(const FamArgs& a, int64_t idx) { p = *reinterpret_cast<const T*>(record(a, idx)); }
    GS_DEV void F(const T (&x)[1], T (&r)[1]) const { r[0] = x[0] * x[0] - p; }
    GS_DEV void J(const T (&x)[1], T (&Jm)[1][1]) const { Jm[0][0] = T(2) * x[0]; }
};

// A x + 0.1 tanh.(x) − b, J = A + Diagonal(0.1 (1 − tanh.(x)²)); record = A (9, column-major) + b (3).
template <class R> struct FamTanhLinear3T {
    using T = R;
    static constexpr int N = 3;
    T A[3][3], b[3];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const T* p = reinterpret_cast<const T*>(record(a, idx));
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) A[i][j] = p[i + 3 * j];
#pragma unroll
        for (int i = 0; i < 3; ++i) b[i] = p[9 + i];
    }
    GS_DEV void F(const T (&x)[3], T (&r)[3]) const {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T ax = A[i][0] * x[0];
            ax = tfma(A[i][1], x[1], ax);
            ax = tfma(A[i][2], x[2], ax);
            r[i] = (ax + T(0.1) * ttanh(x[i])) - b[i];
        }
    }
    GS_DEV void J(const T (&x)[3], T (&Jm)[3][3]) const {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) Jm[i][j] = A[i][j];
            T t = ttanh(x[i]);
            Jm[i][i] = A[i][i] + T(0.1) * (T(1) - t * t);
        }
    }
};

// Implicit midpoint of Hénon–Heiles, x = (q₁, q₂, p₁, p₂), record = x_k (4), shared[0] = h:
// F = x − x_k − h f((x_k + x)/2), f = (p₁, p₂, −q₁ − 2q₁q₂, −q₂ − q₁² + q₂²), J = I − (h/2) Df.
template <class R> struct FamHenonHeilesMidpointT {
    using T = R;
    static constexpr int N = 4;
    T h, xk[4];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        if constexpr (sizeof(T) == 8) {
            const double2* p = reinterpret_cast<const double2*>(record(a, idx));
            double2 u = p[0], v = p[1];
            xk[0] = u.x;
            xk[1] = u.y;
            xk[2] = v.x;
            xk[3] = v.y;
        } else {
            const float4 u = *reinterpret_cast<const float4*>(record(a, idx));
            xk[0] = u.x;
            xk[1] = u.y;
            xk[2] = u.z;
            xk[3] = u.w;
        }
        h = T(a.shared[0]);
    }
    GS_DEV void F(const T (&x)[4], T (&r)[4]) const {
        T m0 = (xk[0] + x[0]) / T(2);
        T m1 = (xk[1] + x[1]) / T(2);
        T m2 = (xk[2] + x[2]) / T(2);
        T m3 = (xk[3] + x[3]) / T(2);
        T f0 = m2;
        T f1 = m3;
        T f2 = -m0 - (T(2) * m0) * m1;
        T f3 = (-m1 - m0 * m0) + m1 * m1;
        r[0] = (x[0] - xk[0]) - h * f0;
        r[1] = (x[1] - xk[1]) - h * f1;
        r[2] = (x[2] - xk[2]) - h * f2;
        r[3] = (x[3] - xk[3]) - h * f3;
    }
    GS_DEV void J(const T (&x)[4], T (&Jm)[4][4]) const {
        T q1 = (xk[0] + x[0]) / T(2);
        T q2 = (xk[1] + x[1]) / T(2);
        T c = h / T(2);
        Jm[0][0] = T(1); Jm[0][1] = T(0); Jm[0][2] = -c;  Jm[0][3] = T(0);
        Jm[1][0] = T(0); Jm[1][1] = T(1); Jm[1][2] = T(0); Jm[1][3] = -c;
        Jm[2][0] = c * (T(1) + T(2) * q2); Jm[2][1] = c * (T(2) * q1); Jm[2][2] = T(1); Jm[2][3] = T(0);
        Jm[3][0] = c * (T(2) * q1); Jm[3][1] = c * (T(1) - T(2) * q2); Jm[3][2] = T(0); Jm[3][3] = T(1);
    }
};

using FamCoupled2 = FamCoupled2T<double>;
using FamHetero2 = FamHetero2T<double>;
using FamScalarSquare = FamScalarSquareT<double>;
using FamTanhLinear3 = FamTanhLinear3T<double>;
using FamHenonHeilesMidpoint = FamHenonHeilesMidpointT<double>;

}  // namespace gs
