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

// _r2_coupled_res / _r2_coupled_jac (test/nonlinear/batched.jl:92-103).
struct FamCoupled2 {
    static constexpr int N = 2;
    double p0, p1;
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const double* p = reinterpret_cast<const double*>(record(a, idx));
        p0 = p[0];
        p1 = p[1];
    }
    GS_DEV void F(const double (&x)[2], double (&r)[2]) const {
        r[0] = (x[0] + (x[1] * x[1]) / 8.0) - p0;
        r[1] = (x[1] + (x[0] * x[0]) / 8.0) - p1;
    }
    GS_DEV void J(const double (&x)[2], double (&Jm)[2][2]) const {
        Jm[0][0] = 1.0;
        Jm[0][1] = x[1] / 4.0;
        Jm[1][0] = x[0] / 4.0;
        Jm[1][1] = 1.0;
    }
};

// _r2_hetero_res / _r2_hetero_jac (test/nonlinear/batched.jl:54-90): one fixture per ReturnCode.
struct FamHetero2 {
    static constexpr int N = 2;
    int64_t code;
    GS_DEV void load(const FamArgs& a, int64_t idx) { code = *reinterpret_cast<const int64_t*>(record(a, idx)); }
    GS_DEV void F(const double (&x)[2], double (&r)[2]) const {
        if (code == 1) {
            r[0] = x[0] - 2.0;
            r[1] = x[1] - 3.0;
        } else if (code == 2 || code == 4) {
            r[0] = 1.0;
            r[1] = 1.0;
        } else if (code == 3) {
            r[0] = x[0] - 100.0;
            r[1] = x[1] - 100.0;
        } else if (code == 5) {
            r[0] = dnan();
            r[1] = dnan();
        } else {
            r[0] = x[0] * x[0] + 1.0;
            r[1] = x[1] * x[1] + 1.0;
        }
    }
    GS_DEV void J(const double (&)[2], double (&Jm)[2][2]) const {
        double d = code == 1 ? 1.0 : code == 2 ? 1e16 : code == 3 ? 2.0 : code == 4 ? 0.0 : code == 5 ? 1.0 : -1.0;
        Jm[0][0] = d;
        Jm[0][1] = 0.0;
        Jm[1][0] = 0.0;
        Jm[1][1] = d;
    }
};

// _r2_scalar_res / _r2_scalar_jac (test/nonlinear/batched.jl:107-108): x² − p, 2x.
struct FamScalarSquare {
    static constexpr int N = 1;
    double p;
    GS_DEV void load(const FamArgs& a, int64_t idx) { p = *reinterpret_cast<const double*>(record(a, idx)); }
    GS_DEV void F(const double (&x)[1], double (&r)[1]) const { r[0] = x[0] * x[0] - p; }
    GS_DEV void J(const double (&x)[1], double (&Jm)[1][1]) const { Jm[0][0] = 2.0 * x[0]; }
};

// A x + 0.1 tanh.(x) − b, J = A + Diagonal(0.1 (1 − tanh.(x)²)); record = A (9, column-major) + b (3).
struct FamTanhLinear3 {
    static constexpr int N = 3;
    double A[3][3], b[3];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const double* p = reinterpret_cast<const double*>(record(a, idx));
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) A[i][j] = p[i + 3 * j];
#pragma unroll
        for (int i = 0; i < 3; ++i) b[i] = p[9 + i];
    }
    GS_DEV void F(const double (&x)[3], double (&r)[3]) const {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double ax = A[i][0] * x[0];
            ax = __builtin_fma(A[i][1], x[1], ax);
            ax = __builtin_fma(A[i][2], x[2], ax);
            r[i] = (ax + 0.1 * tanh(x[i])) - b[i];
        }
    }
    GS_DEV void J(const double (&x)[3], double (&Jm)[3][3]) const {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) Jm[i][j] = A[i][j];
            double t = tanh(x[i]);
            Jm[i][i] = A[i][i] + 0.1 * (1.0 - t * t);
        }
    }
};

// Implicit midpoint of Hénon–Heiles, x = (q₁, q₂, p₁, p₂), record = x_k (4), shared[0] = h:
// F = x − x_k − h f((x_k + x)/2), f = (p₁, p₂, −q₁ − 2q₁q₂, −q₂ − q₁² + q₂²), J = I − (h/2) Df.
struct FamHenonHeilesMidpoint {
    static constexpr int N = 4;
    double h, xk[4];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const double2* p = reinterpret_cast<const double2*>(record(a, idx));
        double2 u = p[0], v = p[1];
        xk[0] = u.x;
        xk[1] = u.y;
        xk[2] = v.x;
        xk[3] = v.y;
        h = a.shared[0];
    }
    GS_DEV void F(const double (&x)[4], double (&r)[4]) const {
        double m0 = (xk[0] + x[0]) / 2.0;
        double m1 = (xk[1] + x[1]) / 2.0;
        double m2 = (xk[2] + x[2]) / 2.0;
        double m3 = (xk[3] + x[3]) / 2.0;
        double f0 = m2;
        double f1 = m3;
        double f2 = -m0 - (2.0 * m0) * m1;
        double f3 = (-m1 - m0 * m0) + m1 * m1;
        r[0] = (x[0] - xk[0]) - h * f0;
        r[1] = (x[1] - xk[1]) - h * f1;
        r[2] = (x[2] - xk[2]) - h * f2;
        r[3] = (x[3] - xk[3]) - h * f3;
    }
    GS_DEV void J(const double (&x)[4], double (&Jm)[4][4]) const {
        double q1 = (xk[0] + x[0]) / 2.0;
        double q2 = (xk[1] + x[1]) / 2.0;
        double c = h / 2.0;
        Jm[0][0] = 1.0; Jm[0][1] = 0.0; Jm[0][2] = -c;  Jm[0][3] = 0.0;
        Jm[1][0] = 0.0; Jm[1][1] = 1.0; Jm[1][2] = 0.0; Jm[1][3] = -c;
        Jm[2][0] = c * (1.0 + 2.0 * q2); Jm[2][1] = c * (2.0 * q1); Jm[2][2] = 1.0; Jm[2][3] = 0.0;
        Jm[3][0] = c * (2.0 * q1); Jm[3][1] = c * (1.0 - 2.0 * q2); Jm[3][2] = 0.0; Jm[3][3] = 1.0;
    }
};

}  // namespace gs
