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

// The index sets and datum of a StationarityProblem batch (gs_stationarity), read by
// FamStationarityT only: E and U 0-based, the datum none, shared (|E| values) or one record per system.
struct StatArgs {
    int32_t e[8];
    int32_t u[8];
    int32_t b_kind;        // 0 none, 1 shared, 2 per system
    int32_t pad_;
    double b_shared[8];
    const void* b_data;    // per-system datum records (device), |E| entries of the working precision
    int64_t b_stride;      // bytes between datum records
};

// The chart of a Reparametrized batch (gs_chart), read by FamReparamT only.
struct ChartArgs {
    int32_t kind;      // 1 linear, 2 polar
    int32_t pad_;
    double B[16];      // linear: B (n × n, row-major)
    double c[4];       // linear: c
};

struct FamArgs {
    const char* P;       // parameter records (device)
    int64_t p_stride;    // bytes between records; 0 = one shared record
    double shared[8];    // scalar parameters shared by every system
    const void* data;    // the family's own device data (gs_family.data: A of the dense family)
    DeflArgs defl;       // the deflation of gs_solve_batch_deflated_f64
    StatArgs stat;       // the index sets and datum of gs_solve_batch_stationarity_f64
    ChartArgs chart;     // the chart of gs_solve_batch_reparametrized_f64
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

// The quadratic map of the reference's mixed-precision tests (test/nonlinear/refinement.jl:40-96):
// u = A x (StaticArrays' muladd fold), F_i = (u_i + c_i u_i²) − b_i, J = Diagonal(1 .+ 2c .* u) · A;
// record = vec(A) (column-major, 4), b (2), c (2). f_mod_r1 and f_ill_r1 take c = (0.1, −0.2).
template <class R> struct FamQuadMap2T {
    using T = R;
    static constexpr int N = 2;
    T A[2][2], b[2], c[2];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const T* p = reinterpret_cast<const T*>(record(a, idx));
        A[0][0] = p[0];
        A[1][0] = p[1];
        A[0][1] = p[2];
        A[1][1] = p[3];
        b[0] = p[4];
        b[1] = p[5];
        c[0] = p[6];
        c[1] = p[7];
    }
    GS_DEV void u_of(const T (&x)[2], T (&u)[2]) const {
#pragma unroll
        for (int i = 0; i < 2; ++i) u[i] = tfma(A[i][1], x[1], A[i][0] * x[0]);
    }
    GS_DEV void F(const T (&x)[2], T (&r)[2]) const {
        T u[2];
        u_of(x, u);
#pragma unroll
        for (int i = 0; i < 2; ++i) r[i] = (u[i] + c[i] * (u[i] * u[i])) - b[i];
    }
    GS_DEV void J(const T (&x)[2], T (&Jm)[2][2]) const {
        T u[2];
        u_of(x, u);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const T d = T(1) + (T(2) * c[i]) * u[i];
#pragma unroll
            for (int j = 0; j < 2; ++j) Jm[i][j] = d * A[i][j];
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

// The sine and cosine of the dense family (family 10), one fixed algorithm that the oracle evaluates
// the same way, so that its F and J have the same bits on both sides: k = rint(x · 2/π); r = x − k·π/2
// with π/2 in three parts (Cody–Waite, each product folded by FMA); the Taylor polynomials of sin r
// and cos r in r² by Horner with FMA (through r¹⁷ and r¹⁶ in FP64, r¹¹ and r¹⁰ in FP32; the
// coefficients (−1)^j/(2j+1)! and (−1)^j/(2j)! rounded once); the quadrant k mod 4. NaN for a
// non-finite x or |x| ≥ 2^50 (2^22 in FP32); sin(±0) = ±0.
GS_DEV void fsincos(double x, double& sn, double& cs) {
    constexpr double kS[8] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
                              0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                              -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49};
    constexpr double kC[8] = {-0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
                              0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
                              -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45};
    if (!(__builtin_fabs(x) < 0x1p50)) {
        sn = cs = dnan();
        return;
    }
    if (x == 0.0) {  // sin(±0) = ±0
        sn = x;
        cs = 1.0;
        return;
    }
    const double k = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, x);
    r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);
    r = __builtin_fma(-k, -0x1.f1976b7ed8fbcp-110, r);
    const double z = r * r;
    double ps = kS[7], pc = kC[7];
#pragma unroll
    for (int j = 6; j >= 0; --j) {
        ps = __builtin_fma(ps, z, kS[j]);
        pc = __builtin_fma(pc, z, kC[j]);
    }
    const double s0 = __builtin_fma(r * z, ps, r);
    const double c0 = __builtin_fma(z, pc, 1.0);
    const int q = int(int64_t(k) & 3);
    sn = q == 0 ? s0 : (q == 1 ? c0 : (q == 2 ? -s0 : -c0));
    cs = q == 0 ? c0 : (q == 1 ? -s0 : (q == 2 ? -c0 : s0));
}

GS_DEV void fsincos(float x, float& sn, float& cs) {
    constexpr float kS[5] = {-0x1.555556p-3f, 0x1.111112p-7f, -0x1.a01a02p-13f, 0x1.71de3ap-19f, -0x1.ae6456p-26f};
    constexpr float kC[5] = {-0x1.000000p-1f, 0x1.555556p-5f, -0x1.6c16c2p-10f, 0x1.a01a02p-16f, -0x1.27e4fcp-22f};
    if (!(__builtin_fabsf(x) < 0x1p22f)) {
        sn = cs = tnan<float>();
        return;
    }
    if (x == 0.0f) {  // sin(±0) = ±0
        sn = x;
        cs = 1.0f;
        return;
    }
    const float k = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float z = r * r;
    float ps = kS[4], pc = kC[4];
#pragma unroll
    for (int j = 3; j >= 0; --j) {
        ps = __builtin_fmaf(ps, z, kS[j]);
        pc = __builtin_fmaf(pc, z, kC[j]);
    }
    const float s0 = __builtin_fmaf(r * z, ps, r);
    const float c0 = __builtin_fmaf(z, pc, 1.0f);
    const int q = int(int64_t(k) & 3);
    sn = q == 0 ? s0 : (q == 1 ? c0 : (q == 2 ? -s0 : -c0));
    cs = q == 0 ? c0 : (q == 1 ? -s0 : (q == 2 ? -c0 : s0));
}

// The dense family (TenEResidual, device_problems.jl:147-157) at n ≤ 8 for the lane-per-system
// kernels: F = A x + sin.(x)/10 − b and J = A + Diagonal(cos.(x)/10), the product as StaticArrays'
// muladd fold over each row. A (column-major, shared by the batch) stays in device memory: every
// lane reads the same entries, so they come through the scalar cache; b is the system's record.
template <class R, int NN> struct FamDenseSinSmallT {
    using T = R;
    static constexpr int N = NN;
    const T* A;
    T b[N];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        A = static_cast<const T*>(a.data);
        const T* p = reinterpret_cast<const T*>(record(a, idx));
#pragma unroll
        for (int i = 0; i < N; ++i) b[i] = p[i];
    }
    GS_DEV void F(const T (&x)[N], T (&r)[N]) const {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = A[i] * x[0];
#pragma unroll
            for (int j = 1; j < N; ++j) acc = tfma(A[i + j * N], x[j], acc);
            T sn, cs;
            fsincos(x[i], sn, cs);
            r[i] = (acc + sn / T(10)) - b[i];
        }
    }
    GS_DEV void J(const T (&x)[N], T (&Jm)[N][N]) const {
#pragma unroll
        for (int i = 0; i < N; ++i) {
#pragma unroll
            for (int j = 0; j < N; ++j) Jm[i][j] = A[i + j * N];
            T sn, cs;
            fsincos(x[i], sn, cs);
            Jm[i][i] = A[i + i * N] + cs / T(10);
        }
    }
};

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

// The discrete action of the variational midpoint rule for Hénon–Heiles over two steps, on the full
// state x = (q₋, q, q₊) (6 entries): S(x) = L_d(q₋, q) + L_d(q, q₊), L_d(a, b) = |b − a|²/(2h) −
// h V((a + b)/2), V(q) = (q₁² + q₂²)/2 + λ (q₁² q₂ − q₂³/3) (λ = 1: the Hénon–Heiles potential of
// the midpoint family). A registered action of StationarityProblem(S; E, U, b): `grad` is ∇S and
// `hess` ∇²S, in closed form. The record is (h, λ).
template <class R> struct ActHenonHeilesT {
    using T = R;
    static constexpr int M = 6;
    T h, lam;
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        const T* p = reinterpret_cast<const T*>(record(a, idx));
        h = p[0];
        lam = p[1];
    }
    // ∇V(m)
    GS_DEV void dV(const T (&m)[2], T (&d)[2]) const {
        const T l2 = T(2) * lam;
        d[0] = m[0] + l2 * (m[0] * m[1]);
        d[1] = m[1] + lam * ((m[0] * m[0]) - (m[1] * m[1]));
    }
    GS_DEV void grad(const T (&x)[6], T (&g)[6]) const {
        const T hh = h / T(2);
        T m1[2], m2[2], d1[2], d2[2], v1[2], v2[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            v1[i] = x[2 + i] - x[i];
            v2[i] = x[4 + i] - x[2 + i];
            m1[i] = (x[i] + x[2 + i]) / T(2);
            m2[i] = (x[2 + i] + x[4 + i]) / T(2);
        }
        dV(m1, d1);
        dV(m2, d2);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            g[i] = -(v1[i] / h) - hh * d1[i];
            g[2 + i] = ((v1[i] / h) - hh * d1[i]) - ((v2[i] / h) + hh * d2[i]);
            g[4 + i] = (v2[i] / h) - hh * d2[i];
        }
    }
    GS_DEV void hess(const T (&x)[6], T (&H)[6][6]) const {
        const T ih = T(1) / h, hq = h / T(4), l2 = T(2) * lam;
        T w[2][3];  // ∇²V at the two midpoints: (w00, w01, w11)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const T m0 = (x[2 * k] + x[2 * k + 2]) / T(2);
            const T m1 = (x[2 * k + 1] + x[2 * k + 3]) / T(2);
            w[k][0] = T(1) + l2 * m1;
            w[k][1] = l2 * m0;
            w[k][2] = T(1) - l2 * m1;
        }
        // P = I/h − (h/4) W, Q = −I/h − (h/4) W, per step k
        T P[2][3], Q[2][3];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            P[k][0] = ih - hq * w[k][0];
            P[k][1] = -(hq * w[k][1]);
            P[k][2] = ih - hq * w[k][2];
            Q[k][0] = -ih - hq * w[k][0];
            Q[k][1] = -(hq * w[k][1]);
            Q[k][2] = -ih - hq * w[k][2];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) H[i][j] = T(0);
        auto put = [&](int bi, int bj, T a00, T a01, T a11) {
            H[2 * bi][2 * bj] = a00;
            H[2 * bi][2 * bj + 1] = a01;
            H[2 * bi + 1][2 * bj] = a01;
            H[2 * bi + 1][2 * bj + 1] = a11;
        };
        put(0, 0, P[0][0], P[0][1], P[0][2]);
        put(0, 1, Q[0][0], Q[0][1], Q[0][2]);
        put(1, 0, Q[0][0], Q[0][1], Q[0][2]);
        put(1, 1, P[0][0] + P[1][0], P[0][1] + P[1][1], P[0][2] + P[1][2]);
        put(1, 2, Q[1][0], Q[1][1], Q[1][2]);
        put(2, 1, Q[1][0], Q[1][1], Q[1][2]);
        put(2, 2, P[1][0], P[1][1], P[1][2]);
    }
};

// StationarityProblem(S; E, U, b) (src/nonlinear/stationarity.jl:1-60) on the unknowns u = x[U] of
// the full state x of an action Act: F_a(u) = (∇S)(x(u))_{E_a} (+ b_a), J_ac = (∇²S)(x(u))_{E_a, U_c},
// with x(u)_i = u_c for i = U_c and the start's x_i elsewhere (_embed_unknowns, :209-226); NE = |E|.
// The kernel loads the full state (`gather`) and writes back only the entries of U (`scatter`), so
// the others keep their bits. Indices are runtime values: every pick is a chain of selects, so no
// register array is indexed dynamically.
template <class Act, int NE> struct FamStationarityT {
    using T = typename Act::T;
    static constexpr int N = NE;
    static constexpr int kStateLen = Act::M;
    static constexpr int M = Act::M;
    Act act;
    T xref[M];
    // em[k][i]: E_k = i; um[c][i]: U_c = i. Held as flags, not indices: a chain of selects on
    // `index == i` is rewritten by the compiler into an indexed stack array (scratch memory).
    bool em[NE][M], um[NE][M];
    bool has_b;
    T b[NE];
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        act.load(a, idx);
#pragma unroll
        for (int k = 0; k < NE; ++k)
#pragma unroll
            for (int i = 0; i < M; ++i) {
                em[k][i] = a.stat.e[k] == i;
                um[k][i] = a.stat.u[k] == i;
            }
        has_b = a.stat.b_kind != 0;
        if (a.stat.b_kind == 2) {
            const T* bp = reinterpret_cast<const T*>(reinterpret_cast<const char*>(a.stat.b_data) + idx * a.stat.b_stride);
#pragma unroll
            for (int k = 0; k < NE; ++k) b[k] = bp[k];
        } else {
            // the kernel argument read at constant offsets (a pointer into it would copy it to scratch)
#pragma unroll
            for (int k = 0; k < NE; ++k) b[k] = has_b ? T(a.stat.b_shared[k]) : T(0);
        }
    }
    GS_DEV static T pick(const T (&v)[M], const bool (&m)[M]) {
        T r = v[0];
#pragma unroll
        for (int i = 1; i < M; ++i) r = m[i] ? v[i] : r;
        return r;
    }
    GS_DEV void gather(const T* xp, T (&uu)[NE]) {
#pragma unroll
        for (int i = 0; i < M; ++i) xref[i] = xp[i];
#pragma unroll
        for (int c = 0; c < NE; ++c) uu[c] = pick(xref, um[c]);
    }
    GS_DEV void scatter(T* xp, const T (&uu)[NE]) const {
        T x[M];
        embed(uu, x);
#pragma unroll
        for (int i = 0; i < M; ++i) xp[i] = x[i];
    }
    GS_DEV void embed(const T (&uu)[NE], T (&x)[M]) const {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            T xi = xref[i];
#pragma unroll
            for (int c = 0; c < NE; ++c) xi = um[c][i] ? uu[c] : xi;
            x[i] = xi;
        }
    }
    GS_DEV void F(const T (&uu)[NE], T (&r)[NE]) const {
        T x[M], g[M];
        embed(uu, x);
        act.grad(x, g);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const T gk = pick(g, em[k]);
            r[k] = has_b ? gk + b[k] : gk;
        }
    }
    GS_DEV void J(const T (&uu)[NE], T (&Jm)[NE][NE]) const {
        T x[M], H[M][M];
        embed(uu, x);
        act.hess(x, H);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            T row[M];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                T r = H[0][j];
#pragma unroll
                for (int i = 1; i < M; ++i) r = em[k][i] ? H[i][j] : r;
                row[j] = r;
            }
#pragma unroll
            for (int c = 0; c < NE; ++c) Jm[k][c] = pick(row, um[c]);
        }
    }
};

// The registered charts x = g(y) of Reparametrized, with their Jacobians G = ∂g/∂y.
// Linear: x = B y + c (each row a left fold), G = B; in the working precision R (a Float32 batch
// takes B and c rounded to Float32, as ToReal{Float32} adapts a chart's captured constants).
template <int N, class R = double> struct ChartLinearT {
    using T = R;
    T B[N][N], c[N];
    GS_DEV void load(const FamArgs& a) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            c[i] = T(a.chart.c[i]);
#pragma unroll
            for (int j = 0; j < N; ++j) B[i][j] = T(a.chart.B[i * N + j]);
        }
    }
    GS_DEV void g(const T (&y)[N], T (&x)[N]) const {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T acc = B[i][0] * y[0];
#pragma unroll
            for (int j = 1; j < N; ++j) acc = acc + B[i][j] * y[j];
            x[i] = acc + c[i];
        }
    }
    GS_DEV void G(const T (&)[N], T (&Gm)[N][N]) const {
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) Gm[i][j] = B[i][j];
    }
};

// Polar (the chart of test/nonlinear/reparametrized.jl:31-38): x = (r cos θ, r sin θ),
// G = [cos θ, −r sin θ; sin θ, r cos θ].
struct ChartPolarT {
    using T = double;
    GS_DEV void load(const FamArgs&) {}
    GS_DEV void g(const double (&y)[2], double (&x)[2]) const {
        x[0] = y[0] * cos(y[1]);
        x[1] = y[0] * sin(y[1]);
    }
    GS_DEV void G(const double (&y)[2], double (&Gm)[2][2]) const {
        const double s = sin(y[1]), c = cos(y[1]);
        Gm[0][0] = c;
        Gm[0][1] = -y[0] * s;
        Gm[1][0] = s;
        Gm[1][1] = y[0] * c;
    }
};

// Reparametrized(NonlinearProblem(F), g) (src/nonlinear/reparametrized.jl:1-60, covariant = false):
// the solve iterates the chart coordinates y with F̃(y) = F(g(y)) and J̃(y) = J(g(y)) G(y) (each
// entry a left fold over k), and the batch receives x* = g(y*) (:860-865). X holds y₀ on entry.
template <class Inner, class Chart> struct FamReparamT {
    using T = typename Inner::T;
    static constexpr int N = Inner::N;
    static constexpr bool kMapsOutput = true;
    static_assert(std::is_same<typename Chart::T, T>::value, "the chart computes in the family's precision");
    Inner in;
    Chart ch;
    GS_DEV void load(const FamArgs& a, int64_t idx) {
        in.load(a, idx);
        ch.load(a);
    }
    GS_DEV void F(const T (&y)[N], T (&r)[N]) const {
        T x[N];
        ch.g(y, x);
        in.F(x, r);
    }
    GS_DEV void J(const T (&y)[N], T (&Jm)[N][N]) const {
        T x[N], Ji[N][N], Gm[N][N];
        ch.g(y, x);
        in.J(x, Ji);
        ch.G(y, Gm);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                T acc = Ji[i][0] * Gm[0][j];
#pragma unroll
                for (int k = 1; k < N; ++k) acc = acc + Ji[i][k] * Gm[k][j];
                Jm[i][j] = acc;
            }
    }
    GS_DEV void output(const T (&y)[N], T (&x)[N]) const { ch.g(y, x); }
};

template <class F, class = void> struct maps_output : std::false_type {};
template <class F> struct maps_output<F, std::void_t<decltype(F::kMapsOutput)>> : std::true_type {};

// Whether a family holds a full state beyond its unknowns (X rows of kStateLen entries).
template <class F, class = void> struct holds_state : std::false_type {};
template <class F> struct holds_state<F, std::void_t<decltype(F::kStateLen)>> : std::true_type {};

// Whether a family is a Deflated wrapper (its status reports the inner residual).
template <class F, class = void> struct is_deflated : std::false_type {};
template <class F> struct is_deflated<F, std::void_t<decltype(F::kDeflated)>> : std::true_type {};

}  // namespace gs
