// The registered problem families as device functions: residual F(x, p) and analytic Jacobian
// J(x, p) of one system, with its parameter record loaded into registers once per system. The Julia
// extension (ext/GeometricSolversHIPExt.jl) defines the same F and J as the `NonlinearProblem` of
// each family, and these follow the operation order of those Julia expressions (no contraction;
// `muladd` only where StaticArrays' matrix-vector product uses it).
//
// A family provides: N (unknowns), `load(args, idx)`, `F(x, r)` and `J(x, Jm)` with Jm[i][j] the
// entry (i, j).
#pragma once

#include <type_traits>

#include "device_common.hpp"

namespace gs {

// Launch arguments shared by every family.
// The roots, power and shift of a Deflated batch (gs_deflation), read by FamDeflatedT only.
struct DeflArgs {
    int32_t nroots;      // 1..4
    int32_t power;       // p, an integer ≥ 1
    double shift;        // σ
    double roots[32];    // root r at roots[r·N ..]
};

struct FamArgs {
    const char* P;       // parameter records (device)
    int64_t p_stride;    // bytes between records; 0 = one shared record
    double shared[8];    // scalar parameters shared by every system
    DeflArgs defl;       // the deflation of gs_solve_batch_deflated_f64
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

// ---- Deflated (src/nonlinear/deflation.jl) --------------------------------------------------------

// Julia's x^n for a Float64 base and an integer n (Base.Math.pow_body, Julia 1.13; not under the
// reference tree), which x^y with an integer-valued Float64 y runs: power by squaring of the
// reciprocal (for n < 0) with a compensated running product, x = 1 giving 1 exactly.
GS_DEV double pow_int(double x, int n) {
    if (x == 1.0 || n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    if (n < 0) {
        const double rx = 1.0 / x;
        if (n == -1) return rx;
        if (isfin(x)) xnlo = -__builtin_fma(x, rx, -1.0) * rx;
        x = rx;
        n = -n;
    }
    while (n > 1) {
        if (n & 1) {
            const double err = __builtin_fma(y, xnlo, x * ynlo);
            const double hi = x * y;
            ynlo = __builtin_fma(x, y, -hi) + err;
            y = hi;
        }
        const double err = (x * 2.0) * xnlo;
        const double hi = x * x;
        xnlo = __builtin_fma(x, x, -hi) + err;
        x = hi;
        n >>= 1;
    }
    const double err = __builtin_fma(y, xnlo, x * ynlo);
    return (isfin(x) && isfin(err)) ? __builtin_fma(x, y, err) : x * y;
}

// Deflated(prob, roots; power = p, shift = σ) around family Inner: the residual M(x) F(x) with
// M = ∏ᵢ (‖x − rᵢ‖^(−p) + σ) (_deflate_scale, deflation.jl:251-261), the Jacobian M J + F ∇Mᵀ
// (_deflated_jacobian, :339-343) with (M, ∇M) by the division-free recurrence g ← mᵢ g + M ∇mᵢ,
// M ← M mᵢ and the unit-scaled coordinates of ∇mᵢ (_deflate_factors, :264-278; :239-241), and the
// product M (J v) + ⟨∇M, v⟩ F (_deflated_jvp, :352-357). Roots beyond nroots are never read.
template <class Inner> struct FamDeflatedT {
    using T = typename Inner::T;
    static constexpr int N = Inner::N;
    static constexpr bool kDeflated = true;
    static_assert(sizeof(T) == 8, "the device deflation is Float64");
    Inner in;
    int nroots, power;
    T shift;
    T roots[4][N];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        in.load(a, idx);
        nroots = a.defl.nroots;
        power = a.defl.power;
        shift = T(a.defl.shift);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int i = 0; i < N; ++i) roots[r][i] = T(a.defl.roots[r * N + i]);
    }
    // M(x)
    GS_DEV T scale(const T (&x)[N]) const {
        T M = T(1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r >= nroots) break;
            T u[N];
#pragma unroll
            for (int i = 0; i < N; ++i) u[i] = x[i] - roots[r][i];
            M = M * (pow_int(rnorm<N>(u), -power) + shift);
        }
        return M;
    }
    // (M(x), ∇M(x))
    GS_DEV void factors(const T (&x)[N], T& M, T (&g)[N]) const {
        M = T(1);
#pragma unroll
        for (int i = 0; i < N; ++i) g[i] = T(0);
        const T pw = T(power);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r >= nroots) break;
            T u[N];
#pragma unroll
            for (int i = 0; i < N; ++i) u[i] = x[i] - roots[r][i];
            const T d = rnorm<N>(u);
            const T phi = pow_int(d, -power);
            const T m = phi + shift;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const T gi = u[i] == T(0) ? T(0) : -((pw * ((u[i] / d) * phi)) / d);
                g[i] = m * g[i] + M * gi;
            }
            M = M * m;
        }
    }
    GS_DEV void F(const T (&x)[N], T (&r)[N]) const {
        const T M = scale(x);
        in.F(x, r);
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = M * r[i];
    }
    GS_DEV void J(const T (&x)[N], T (&Jm)[N][N]) const {
        T Fi[N], M, g[N];
        in.F(x, Fi);
        in.J(x, Jm);
        factors(x, M, g);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) Jm[i][j] = M * Jm[i][j] + Fi[i] * g[j];
    }
    GS_DEV void Jv(const T (&x)[N], const T (&v)[N], T (&out)[N]) const {
        T Fi[N], Ji[N][N], jv[N], M, g[N];
        in.F(x, Fi);
        in.J(x, Ji);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = Ji[i][0] * v[0];
#pragma unroll
            for (int k = 1; k < N; ++k) acc = __builtin_fma(Ji[i][k], v[k], acc);
            jv[i] = acc;
        }
        factors(x, M, g);
        const T gv = rdot<N>(g, v);
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = M * jv[i] + gv * Fi[i];
    }
};

// Whether a family is a Deflated wrapper (its status reports the inner residual).
template <class F, class = void> struct is_deflated : std::false_type {};
template <class F> struct is_deflated<F, std::void_t<decltype(F::kDeflated)>> : std::true_type {};

}  // namespace gs
