// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// The registered problem families, restated from their Julia definitions (the `F` and `J` of the
// `NonlinearProblem`s that ext/GeometricSolversHIPExt.jl registers, and the fixtures of the
// reference's own batched tests). Each evaluation follows the Julia expression's operation order
// with no contraction; `nf` counts residual evaluations.
#pragma once

#include "oracle.hpp"

namespace orc {

// _r2_coupled_res / _r2_coupled_jac (test/nonlinear/batched.jl:92-103):
//   F = (x₁ + x₂²/8 − p₁, x₂ + x₁²/8 − p₂),  J = [1 x₂/4; x₁/4 1].
template <class T> struct Coupled2 {
    static constexpr int N = 2;
    T p0, p1;
    long nf = 0;
    Vec<T, 2> F(const Vec<T, 2>& x) {
        ++nf;
        return {(x[0] + (x[1] * x[1]) / T(8)) - p0, (x[1] + (x[0] * x[0]) / T(8)) - p1};
    }
    Mat<T, 2> J(const Vec<T, 2>& x) const { return {T(1), x[0] / T(4), x[1] / T(4), T(1)}; }
};

// _r2_hetero_res / _r2_hetero_jac (test/nonlinear/batched.jl:54-90): six fixtures selected by an
// integer code, one per ReturnCode.
template <class T> struct Hetero2 {
    static constexpr int N = 2;
    int64_t code;
    long nf = 0;
    Vec<T, 2> F(const Vec<T, 2>& x) {
        ++nf;
        switch (code) {
            case 1: return {x[0] - T(2), x[1] - T(3)};
            case 2: return {T(1), T(1)};
            case 3: return {x[0] - T(100), x[1] - T(100)};
            case 4: return {T(1), T(1)};
            case 5: return {nan_<T>(), nan_<T>()};
            default: return {x[0] * x[0] + T(1), x[1] * x[1] + T(1)};
        }
    }
    Mat<T, 2> J(const Vec<T, 2>&) const {
        switch (code) {
            case 1: return {T(1), T(0), T(0), T(1)};
            case 2: return {T(1e16), T(0), T(0), T(1e16)};
            case 3: return {T(2), T(0), T(0), T(2)};
            case 4: return {T(0), T(0), T(0), T(0)};
            case 5: return {T(1), T(0), T(0), T(1)};
            default: return {T(-1), T(0), T(0), T(-1)};
        }
    }
};

// _r2_scalar_res / _r2_scalar_jac (test/nonlinear/batched.jl:107-108): F = x² − p, J = 2x.
template <class T> struct ScalarSquare {
    static constexpr int N = 1;
    T p;
    long nf = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {x[0] * x[0] - p};
    }
    Mat<T, 1> J(const Vec<T, 1>& x) const { return {T(2) * x[0]}; }
};

// The n = 3 linear-plus-tanh family: F(x, (A, b)) = A x + 0.1 tanh.(x) − b and
// J = A + Diagonal(0.1 (1 − tanh.(x)²)); A x is StaticArrays' muladd fold.
template <class T> struct TanhLinear3 {
    static constexpr int N = 3;
    Mat<T, 3> A;
    Vec<T, 3> b;
    long nf = 0;
    Vec<T, 3> F(const Vec<T, 3>& x) {
        ++nf;
        Vec<T, 3> Ax = matvec<T, 3>(A, x);
        Vec<T, 3> r;
        for (int i = 0; i < 3; ++i) r[i] = (Ax[i] + T(0.1) * std::tanh(x[i])) - b[i];
        return r;
    }
    Mat<T, 3> J(const Vec<T, 3>& x) const {
        Mat<T, 3> J = A;
        for (int i = 0; i < 3; ++i) {
            T t = std::tanh(x[i]);
            J[i + 3 * i] = A[i + 3 * i] + T(0.1) * (T(1) - t * t);
        }
        return J;
    }
};

// The implicit-midpoint step of the Hénon–Heiles system, x = (q₁, q₂, p₁, p₂):
//   m = (x_k + x)/2,  f(m) = (p₁, p₂, −q₁ − 2q₁q₂, −q₂ − q₁² + q₂²),  F = x − x_k − h f(m),
//   J = I − (h/2) Df(m).
template <class T> struct HenonHeilesMidpoint {
    static constexpr int N = 4;
    T h;
    Vec<T, 4> xk;
    long nf = 0;
    Vec<T, 4> F(const Vec<T, 4>& x) {
        ++nf;
        Vec<T, 4> m;
        for (int i = 0; i < 4; ++i) m[i] = (xk[i] + x[i]) / T(2);
        T f[4] = {m[2], m[3], -m[0] - (T(2) * m[0]) * m[1], (-m[1] - m[0] * m[0]) + m[1] * m[1]};
        Vec<T, 4> r;
        for (int i = 0; i < 4; ++i) r[i] = (x[i] - xk[i]) - h * f[i];
        return r;
    }
    Mat<T, 4> J(const Vec<T, 4>& x) const {
        T q1 = (xk[0] + x[0]) / T(2);
        T q2 = (xk[1] + x[1]) / T(2);
        T c = h / T(2);
        // rows: (1, 0, −c, 0), (0, 1, 0, −c), (c(1+2q₂), c(2q₁), 1, 0), (c(2q₁), c(1−2q₂), 0, 1)
        Mat<T, 4> J;
        auto S = [&](int i, int j, T v) { J[i + 4 * j] = v; };
        S(0, 0, T(1)); S(0, 1, T(0)); S(0, 2, -c); S(0, 3, T(0));
        S(1, 0, T(0)); S(1, 1, T(1)); S(1, 2, T(0)); S(1, 3, -c);
        S(2, 0, c * (T(1) + T(2) * q2)); S(2, 1, c * (T(2) * q1)); S(2, 2, T(1)); S(2, 3, T(0));
        S(3, 0, c * (T(2) * q1)); S(3, 1, c * (T(1) - T(2) * q2)); S(3, 2, T(0)); S(3, 3, T(1));
        return J;
    }
};

// The Gauss–Legendre s = 3 stage system of a 16-site FPU-β chain with fixed ends, n = 48.
//
// The chain: positions q_1..q_16 (q_0 = q_17 = 0), H = Σ p²/2 + Σ_{s=0}^{16} φ(q_{s+1} − q_s),
// φ(r) = r²/2 + β r⁴/4, force g_m = φ'(q_{m+1} − q_m) − φ'(q_m − q_{m−1}), φ'(r) = r + β r³.
// The reference's StageResidual (src/nonlinear/problem.jl:183-189) stacks the first-order stage
// increments of (q, p) (96 unknowns); this family is its position-only (Nyström) form: with
// P_i = p + h Σ_j a_ij g(Q_j) substituted into Q_i = q + h Σ_j a_ij P_j,
//     Q_i = q + h c_i p + h² Σ_j (A²)_ij g(Q_j),
// and the unknowns are the position increments z_i = Q_i − q (start z = 0, the reference's
// increment convention):
//     R_{a,m}(z) = (z_{a,m} − (h c_a) p_m) − h² ((A²_{a0} g_{0,m} + A²_{a1} g_{1,m}) + A²_{a2} g_{2,m}),
//     J_{(a,m),(b,k)} = δ − (h² A²_{ab}) H^b_{mk},  H^b the tridiagonal Hessian of −V at Q_b:
//     H^b_{mm} = −(K_{b,m} + K_{b,m+1}), H^b_{m,m−1} = K_{b,m}, H^b_{m,m+1} = K_{b,m+1},
//     K_{b,s} = φ''(r_{b,s}) = 1 + (3β) r², r_{b,s} = Q_{b,s} − Q_{b,s−1} (bond s, ends at 0).
// c and A² are the Gauss–Legendre tableau (c = ½ ∓ √15/10, ½; A the collocation matrix of
// test/helpers/stage_problems.jl:148-160) computed in 50-digit arithmetic and rounded once.
constexpr double kGlC[3] = {0x1.cda042f0236e1p-4, 0x1.0000000000000p-1, 0x1.c64bf7a1fb924p-1};
constexpr double kGlA2[3][3] = {{0x1.6c16c16c16c17p-7, -0x1.0f9922995a9d1p-7, 0x1.ce741a1dacb5ep-9},
                                {0x1.a370da4233f7dp-4, 0x1.c71c71c71c71cp-6, -0x1.537f6b3fb1445p-8},
                                {0x1.bfe2a15ea5befp-3, 0x1.4f8d7b6829931p-3, 0x1.6c16c16c16c17p-7}};

template <class T> struct FpuGl3 {
    static constexpr int N = 48;
    T h, beta;
    Vec<T, 16> q, p;
    long nf = 0;
    // Q_b,m = q_m + z_{b,m}
    void stages(const Vec<T, 48>& z, T (&Q)[3][16]) const {
        for (int b = 0; b < 3; ++b)
            for (int m = 0; m < 16; ++m) Q[b][m] = q[m] + z[b * 16 + m];
    }
    static T bond(const T (&Qb)[16], int s) {  // r_s = Q_s − Q_{s−1}, the ends at 0
        T hi = s < 16 ? Qb[s] : T(0);
        T lo = s > 0 ? Qb[s - 1] : T(0);
        return hi - lo;
    }
    Vec<T, 48> F(const Vec<T, 48>& z) {
        ++nf;
        T Q[3][16], g[3][16];
        stages(z, Q);
        for (int b = 0; b < 3; ++b)
            for (int m = 0; m < 16; ++m) {
                T rl = bond(Q[b], m), rr = bond(Q[b], m + 1);
                T dl = rl + beta * ((rl * rl) * rl), dr = rr + beta * ((rr * rr) * rr);
                g[b][m] = dr - dl;
            }
        const T h2 = h * h;
        Vec<T, 48> R;
        for (int a = 0; a < 3; ++a) {
            const T hc = h * T(kGlC[a]);
            for (int m = 0; m < 16; ++m) {
                T s = T(kGlA2[a][0]) * g[0][m];
                s = s + T(kGlA2[a][1]) * g[1][m];
                s = s + T(kGlA2[a][2]) * g[2][m];
                R[a * 16 + m] = (z[a * 16 + m] - hc * p[m]) - h2 * s;
            }
        }
        return R;
    }
    Mat<T, 48> J(const Vec<T, 48>& z) const {
        T Q[3][16], K[3][17];
        stages(z, Q);
        for (int b = 0; b < 3; ++b)
            for (int s = 0; s <= 16; ++s) {
                T r = bond(Q[b], s);
                K[b][s] = T(1) + (T(3) * beta) * (r * r);
            }
        const T h2 = h * h;
        Mat<T, 48> J;
        for (int j = 0; j < 48; ++j) {
            const int b = j / 16, k = j % 16;
            for (int i = 0; i < 48; ++i) {
                const int a = i / 16, m = i % 16;
                T H = T(0);
                if (k == m) H = -(K[b][m] + K[b][m + 1]);
                else if (k + 1 == m) H = K[b][m];
                else if (k == m + 1) H = K[b][m + 1];
                const T coef = h2 * T(kGlA2[a][b]);
                J[i + 48 * j] = (i == j ? T(1) : T(0)) - coef * H;
            }
        }
        return J;
    }
};

// Fixtures of the reference's loop tests (test/nonlinear/solver.jl:183-290), for the oracle's KATs:
//   101: y.^2 − q, J = Diagonal(2y), n = 2         (solver.jl:242-246)
//   102: 2y − q,  J = 2I, n = 2                    (solver.jl:279-281)
//   103: x·x − 4, J = Inf, n = 1 (a zero step)     (solver.jl:185)
// `nj` counts the calls of J, the factorisations' and the JVPs' (J(x) v) alike, as the counting
// Jacobian closures of the reference's tests do (test/nonlinear/newton.jl:486-545).
template <class T> struct DiagSquare2 {
    static constexpr int N = 2;
    T q0, q1;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 2> F(const Vec<T, 2>& y) {
        ++nf;
        return {y[0] * y[0] - q0, y[1] * y[1] - q1};
    }
    Mat<T, 2> J(const Vec<T, 2>& y) const {
        ++nj;
        return {y[0] + y[0], T(0), T(0), y[1] + y[1]};
    }
};
template <class T> struct Affine2 {
    static constexpr int N = 2;
    T q0, q1;
    long nf = 0;
    Vec<T, 2> F(const Vec<T, 2>& y) {
        ++nf;
        return {T(2) * y[0] - q0, T(2) * y[1] - q1};
    }
    mutable long nj = 0;
    Mat<T, 2> J(const Vec<T, 2>&) const {
        ++nj;
        return {T(2), T(0), T(0), T(2)};
    }
};
template <class T> struct InfJacobian1 {
    static constexpr int N = 1;
    long nf = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {x[0] * x[0] - T(4)};
    }
    mutable long nj = 0;
    Mat<T, 1> J(const Vec<T, 1>&) const {
        ++nj;
        return {std::numeric_limits<T>::infinity()};
    }
};

// The one-dimensional residuals of the reference's DogLeg tests (test/globalization/dogleg.jl:
// 48-51, 316-340), all with J = 1, so that each test sets the ratio ρ = ared/pred exactly:
//   kind 0 (pw_F):  p0 at x = 0, p1 at any other x
//   kind 1 (F5):    −10 at x = 0, −10.5 at x ≥ 5, x − 10 otherwise
//   kind 2:         x − 10
template <class T> struct DogLegPw {
    static constexpr int N = 1;
    int kind;
    T p0, p1;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        const T v = x[0];
        if (kind == 0) return {v == T(0) ? p0 : p1};
        if (kind == 1) return {v == T(0) ? T(-10) : (v >= T(5) ? T(-10.5) : v - T(10))};
        return {v - T(10)};
    }
    Mat<T, 1> J(const Vec<T, 1>&) const {
        ++nj;
        return {T(1)};
    }
};

// Family 9, the quadratic map of the reference's mixed-precision tests (test/nonlinear/refinement.jl:
// 40-96): u = A x, F_i = (u_i + c_i u_i²) − b_i, J = Diagonal(1 .+ 2c .* u) · A. f_mod_r1 / j_mod_r1
// and f_ill_r1 / j_ill_r1 are A_MOD, B_MOD and A_ILL, B_ILL with c = (0.1, −0.2): u₂ − 0.2 u₂² and
// u₂ + (−0.2) u₂² are the same IEEE operations, and so are T(0.2), T(0.4) and 2 T(0.1), 2 T(0.2).
// Record: A (4, column-major), b (2), c (2).
template <class T> struct QuadMap2 {
    static constexpr int N = 2;
    Mat<T, 2> A;
    Vec<T, 2> b, c;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 2> F(const Vec<T, 2>& x) {
        ++nf;
        const Vec<T, 2> u = matvec<T, 2>(A, x);
        return {(u[0] + c[0] * (u[0] * u[0])) - b[0], (u[1] + c[1] * (u[1] * u[1])) - b[1]};
    }
    Mat<T, 2> J(const Vec<T, 2>& x) const {
        ++nj;
        const Vec<T, 2> u = matvec<T, 2>(A, x);
        const T d0 = T(1) + (T(2) * c[0]) * u[0];
        const T d1 = T(1) + (T(2) * c[1]) * u[1];
        return {d0 * A[0], d1 * A[1], d0 * A[2], d1 * A[3]};
    }
};

// The sine and cosine of family 10, one fixed algorithm that the device family writes out the same
// way, so that F and J have the same bits on both sides (as the exponential of family 7): k =
// rint(x · 2/π); r = x − k·π/2 with π/2 split in three parts (Cody–Waite, each product folded by
// FMA); the Taylor polynomials of sin r and cos r in r² on |r| ≤ π/4 by Horner with FMA (through r¹⁷
// and r¹⁶ in Float64, r¹¹ and r¹⁰ in Float32); then the quadrant k mod 4. The Taylor coefficients
// (−1)^j/(2j+1)! and (−1)^j/(2j)! are rounded once to T. Within 2 ulp of the libm values on the
// tested range (tests/test_oracle_dense.py); NaN for a non-finite x and for |x| ≥ 2^50 (2^22 in
// Float32), where k would not be an exact integer; sin(±0) = ±0.
template <class T> struct SinCosConst;
template <> struct SinCosConst<double> {
    static constexpr int terms = 8;
    static constexpr double two_over_pi = 0x1.45f306dc9c883p-1;
    static constexpr double p1 = 0x1.921fb54442d18p+0, p2 = 0x1.1a62633145c07p-54, p3 = -0x1.f1976b7ed8fbcp-110;
    static constexpr double big = 0x1p50;
    static constexpr double s[8] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
                                    0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                                    -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49};
    static constexpr double c[8] = {-0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
                                    0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
                                    -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45};
};
template <> struct SinCosConst<float> {
    static constexpr int terms = 5;
    static constexpr float two_over_pi = 0x1.45f306p-1f;
    static constexpr float p1 = 0x1.921fb6p+0f, p2 = -0x1.777a5cp-25f, p3 = -0x1.ee59dap-50f;
    static constexpr float big = 0x1p22f;
    static constexpr float s[8] = {-0x1.555556p-3f, 0x1.111112p-7f, -0x1.a01a02p-13f, 0x1.71de3ap-19f,
                                   -0x1.ae6456p-26f, 0.0f, 0.0f, 0.0f};
    static constexpr float c[8] = {-0x1.000000p-1f, 0x1.555556p-5f, -0x1.6c16c2p-10f, 0x1.a01a02p-16f,
                                   -0x1.27e4fcp-22f, 0.0f, 0.0f, 0.0f};
};

template <class T> inline void fsincos(T x, T& sn, T& cs) {
    using K = SinCosConst<T>;
    if (!(std::fabs(x) < K::big)) {
        sn = cs = std::numeric_limits<T>::quiet_NaN();
        return;
    }
    if (x == T(0)) {  // sin(±0) = ±0
        sn = x;
        cs = T(1);
        return;
    }
    const T k = std::rint(x * K::two_over_pi);
    T r = std::fma(-k, K::p1, x);
    r = std::fma(-k, K::p2, r);
    r = std::fma(-k, K::p3, r);
    const T z = r * r;
    T ps = K::s[K::terms - 1], pc = K::c[K::terms - 1];
    for (int j = K::terms - 2; j >= 0; --j) {
        ps = std::fma(ps, z, K::s[j]);
        pc = std::fma(pc, z, K::c[j]);
    }
    const T s0 = std::fma(r * z, ps, r);
    const T c0 = std::fma(z, pc, T(1));
    const int q = int(int64_t(k) & 3);
    sn = q == 0 ? s0 : (q == 1 ? c0 : (q == 2 ? -s0 : -c0));
    cs = q == 0 ? c0 : (q == 1 ? -s0 : (q == 2 ? -c0 : s0));
}

// Family 10, the §10.E problem of the reference's device benchmark (benchmark/device_problems.jl:
// 133-175, TenEResidual): F(x, b) = A x + sin.(x) ./ 10 − b, with A shared by the batch (N × N,
// column-major; the reference keeps it inside the residual) and b per system. J = A +
// Diagonal(cos.(x) ./ 10): ForwardDiff's partials of the same expression (the matrix-vector
// product's partials are A's entries exactly; the sine's are cos(x_i)·1, divided by 10).
template <class T, int NN> struct DenseSin {
    static constexpr int N = NN;
    const T* A;
    Vec<T, NN> b;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, NN> F(const Vec<T, NN>& x) {
        ++nf;
        Vec<T, NN> r;
        for (int i = 0; i < NN; ++i) {
            T acc = A[i] * x[0];
            for (int j = 1; j < NN; ++j) acc = std::fma(A[i + j * NN], x[j], acc);
            T sn, cs;
            fsincos<T>(x[i], sn, cs);
            r[i] = (acc + sn / T(10)) - b[i];
        }
        return r;
    }
    Mat<T, NN> J(const Vec<T, NN>& x) const {
        ++nj;
        Mat<T, NN> J;
        for (int k = 0; k < NN * NN; ++k) J[k] = A[k];
        for (int i = 0; i < NN; ++i) {
            T sn, cs;
            fsincos<T>(x[i], sn, cs);
            J[i + i * NN] = A[i + i * NN] + cs / T(10);
        }
        return J;
    }
};

// Fixtures of the reference's mixed-precision and stopping tests, for the oracle's KATs:
//   104: f_floor_r1 / j_floor_r1 (refinement.jl:131-139): F rounds x through Float32, so ‖F‖ floors
//        near eps(Float32); J = Diagonal(2x) in Float64
//   105: A x + c .* x.^3 − b, J = A + Diagonal(3c .* x.^2) (refinement.jl:690-720, Check 1); p = vec(A)
//        column-major, c, b
//   106: linear_problem(a) with linear_newton(b) (stopping.jl:28-29): F = a x, J = b; p = (a, b)
//   107: F = x with J = −1 at x = 1 and 2 elsewhere (stopping.jl:705-712)
//   108: FloorResidual (stopping.jl:661-666): F = (x − 1) + δ cos(x / (δ 10⁻⁶)), J = 1; p = (δ)
//   109: CycleStep (stopping.jl:553-563) on F = x: the steps 1 → 1/2 → 1/4 → −1/4 → 1/4 → −1/4 …;
//        p = (limited)
//   110: JumpStep (stopping.jl:582-590) on F = x − 100: the steps 0 → 1 → 2 → 2 + 1/1024, then by 1
// 109 and 110 are methods, not residuals: the solve loop takes their `custom_step` in place of the
// Newton step (their J is never read).
struct Floor2 {
    static constexpr int N = 2;
    long nf = 0;
    mutable long nj = 0;
    Vec<double, 2> F(const Vec<double, 2>& x) {
        ++nf;
        const float a = float(x[0]), b = float(x[1]);
        return {double(a * a - 2.0f), double(b * b - 3.0f)};
    }
    Mat<double, 2> J(const Vec<double, 2>& x) const {
        ++nj;
        return {2.0 * x[0], 0.0, 0.0, 2.0 * x[1]};
    }
};
template <class T> struct Cubic2 {
    static constexpr int N = 2;
    Mat<T, 2> A;
    Vec<T, 2> c, b;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 2> F(const Vec<T, 2>& x) {
        ++nf;
        const Vec<T, 2> Ax = matvec<T, 2>(A, x);
        Vec<T, 2> r;
        for (int i = 0; i < 2; ++i) r[i] = (Ax[i] + c[i] * ((x[i] * x[i]) * x[i])) - b[i];
        return r;
    }
    Mat<T, 2> J(const Vec<T, 2>& x) const {
        ++nj;
        Mat<T, 2> J = A;
        for (int i = 0; i < 2; ++i) J[i + 2 * i] = A[i + 2 * i] + (T(3) * c[i]) * (x[i] * x[i]);
        return J;
    }
};
template <class T> struct LinearConstJ1 {
    static constexpr int N = 1;
    T a, b;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {a * x[0]};
    }
    Mat<T, 1> J(const Vec<T, 1>&) const {
        ++nj;
        return {b};
    }
};
template <class T> struct JumpJacobian1 {
    static constexpr int N = 1;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {x[0]};
    }
    Mat<T, 1> J(const Vec<T, 1>& x) const {
        ++nj;
        return {x[0] == T(1) ? T(-1) : T(2)};
    }
};
template <class T> struct FloorResidual1 {
    static constexpr int N = 1;
    T delta;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {(x[0] - T(1)) + delta * std::cos(x[0] / (delta * T(1e-6)))};
    }
    Mat<T, 1> J(const Vec<T, 1>&) const {
        ++nj;
        return {T(1)};
    }
};
template <class T> struct CycleStep1 {
    static constexpr int N = 1;
    bool limited;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {x[0]};
    }
    Mat<T, 1> J(const Vec<T, 1>&) const { return {T(1)}; }
    StepInfo<T> custom_step(Vec<T, 1>& x, Vec<T, 1>& Fx) {
        const T v = x[0];
        const T h = T(1) / T(2), q = T(1) / T(4);
        const T vp = v == T(1) ? h : (v == h ? q : (v == q ? -q : q));
        const Vec<T, 1> xp{vp};
        const Vec<T, 1> Fp = F(xp);
        StepInfo<T> info = step_info<T>(SUCCESS, std::fabs(Fp[0]), std::fabs(vp - v), 0, false, limited);
        x = xp;
        Fx = Fp;
        return info;
    }
};
template <class T> struct JumpStep1 {
    static constexpr int N = 1;
    long nf = 0;
    mutable long nj = 0;
    Vec<T, 1> F(const Vec<T, 1>& x) {
        ++nf;
        return {x[0] - T(100)};
    }
    Mat<T, 1> J(const Vec<T, 1>&) const { return {T(1)}; }
    StepInfo<T> custom_step(Vec<T, 1>& x, Vec<T, 1>& Fx) {
        const T v = x[0];
        const T vp = v == T(0) ? T(1) : (v == T(1) ? T(2) : (v == T(2) ? T(2) + T(1) / T(1024) : v + T(1)));
        const Vec<T, 1> xp{vp};
        const Vec<T, 1> Fp = F(xp);
        StepInfo<T> info = step_info<T>(SUCCESS, std::fabs(Fp[0]), std::fabs(vp - v), 0, false, false);
        x = xp;
        Fx = Fp;
        return info;
    }
};

// The two-step discrete action of the variational midpoint rule for Hénon–Heiles, a registered
// action of StationarityProblem: x = (q₋, q, q₊) (6 entries), S(x) = L_d(q₋, q) + L_d(q, q₊),
// L_d(a, b) = |b − a|²/(2h) − h V((a + b)/2), V(q) = (q₁² + q₂²)/2 + λ (q₁² q₂ − q₂³/3). ∇S and ∇²S
// in closed form (the reference differentiates S by nested ForwardDiff duals, hessian.jl:74-99; the
// closed forms agree with those to rounding, and the device evaluates these same expressions).
template <class T> struct HenonHeilesAction {
    static constexpr int M = 6;
    T h, lam;
    void dV(const T (&m)[2], T (&d)[2]) const {
        const T l2 = T(2) * lam;
        d[0] = m[0] + l2 * (m[0] * m[1]);
        d[1] = m[1] + lam * ((m[0] * m[0]) - (m[1] * m[1]));
    }
    Vec<T, 6> grad(const Vec<T, 6>& x) const {
        const T hh = h / T(2);
        T m1[2], m2[2], d1[2], d2[2], v1[2], v2[2];
        for (int i = 0; i < 2; ++i) {
            v1[i] = x[2 + i] - x[i];
            v2[i] = x[4 + i] - x[2 + i];
            m1[i] = (x[i] + x[2 + i]) / T(2);
            m2[i] = (x[2 + i] + x[4 + i]) / T(2);
        }
        dV(m1, d1);
        dV(m2, d2);
        Vec<T, 6> g;
        for (int i = 0; i < 2; ++i) {
            g[i] = -(v1[i] / h) - hh * d1[i];
            g[2 + i] = ((v1[i] / h) - hh * d1[i]) - ((v2[i] / h) + hh * d2[i]);
            g[4 + i] = (v2[i] / h) - hh * d2[i];
        }
        return g;
    }
    // H[i][j] = ∂²S/∂xᵢ∂xⱼ: blocks P = I/h − (h/4) W and Q = −I/h − (h/4) W of each step, W = ∇²V
    // at the step's midpoint; H = [P₁ Q₁ 0; Q₁ P₁+P₂ Q₂; 0 Q₂ P₂].
    void hess(const Vec<T, 6>& x, T (&H)[6][6]) const {
        const T ih = T(1) / h, hq = h / T(4), l2 = T(2) * lam;
        T P[2][3], Q[2][3];
        for (int k = 0; k < 2; ++k) {
            const T m0 = (x[2 * k] + x[2 * k + 2]) / T(2);
            const T m1 = (x[2 * k + 1] + x[2 * k + 3]) / T(2);
            const T w0 = T(1) + l2 * m1, w1 = l2 * m0, w2 = T(1) - l2 * m1;
            P[k][0] = ih - hq * w0;
            P[k][1] = -(hq * w1);
            P[k][2] = ih - hq * w2;
            Q[k][0] = -ih - hq * w0;
            Q[k][1] = -(hq * w1);
            Q[k][2] = -ih - hq * w2;
        }
        for (int i = 0; i < 6; ++i)
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

// The reduced problem of a StationarityProblem (_ReducedStationarityProblem, stationarity.jl:305-
// 324) over the action A: the unknowns u = x[U] of the start x₀ (_extract_unknowns, :209-212), the
// full state x(u) with x_{U_c} = u_c and x₀ elsewhere (_embed_unknowns, :221-226), the residual
// (∇S)_E + b, b added only where a datum is given (_stationarity_residual_r1, hessian.jl:47-63),
// and the Jacobian (∇²S)_{E,U} (hessian.jl:74-99). e and u are 0-based.
template <class R, int N, class A> struct StationarityProb {
    static constexpr int M = A::M;
    A act;
    Vec<R, M> xref;
    int e[N], u[N];
    bool has_b = false;
    Vec<R, N> b{};
    long nf = 0;
    Vec<R, M> embed(const Vec<R, N>& uu) const {
        Vec<R, M> x = xref;
        for (int c = 0; c < N; ++c) x[u[c]] = uu[c];
        return x;
    }
    Vec<R, N> extract(const Vec<R, M>& x) const {
        Vec<R, N> uu;
        for (int c = 0; c < N; ++c) uu[c] = x[u[c]];
        return uu;
    }
    Vec<R, N> F(const Vec<R, N>& uu) {
        ++nf;
        const Vec<R, M> g = act.grad(embed(uu));
        Vec<R, N> f;
        for (int a = 0; a < N; ++a) f[a] = has_b ? g[e[a]] + b[a] : g[e[a]];
        return f;
    }
    Mat<R, N> J(const Vec<R, N>& uu) const {
        R H[M][M];
        act.hess(embed(uu), H);
        Mat<R, N> Jm;
        for (int a = 0; a < N; ++a)
            for (int c = 0; c < N; ++c) Jm[a + c * N] = H[e[a]][u[c]];
        return Jm;
    }
};

// The registered charts of Reparametrized (x = g(y), G = ∂g/∂y) and the reparametrized problem
// F̃(y) = F(g(y)), J̃ = J(g(y)) G(y) (_InnerReparametrized, reparametrized.jl:1-60 with
// covariant = false; each entry of J G a left fold over k).
// The linear chart in the working precision R_ (a Float32 batch holds B and c rounded to Float32, as
// ToReal{Float32} adapts a chart's captured constants).
template <int N, class R_ = double> struct ChartLinear {
    using R = R_;
    R B[N][N], c[N];
    Vec<R, N> g(const Vec<R, N>& y) const {
        Vec<R, N> x;
        for (int i = 0; i < N; ++i) {
            R acc = B[i][0] * y[0];
            for (int j = 1; j < N; ++j) acc = acc + B[i][j] * y[j];
            x[i] = acc + c[i];
        }
        return x;
    }
    void G(const Vec<R, N>&, R (&Gm)[N][N]) const {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) Gm[i][j] = B[i][j];
    }
};

// g_polar / dg_polar of test/nonlinear/reparametrized.jl:31-38.
struct ChartPolar {
    using R = double;
    Vec<double, 2> g(const Vec<double, 2>& y) const { return {y[0] * std::cos(y[1]), y[0] * std::sin(y[1])}; }
    void G(const Vec<double, 2>& y, double (&Gm)[2][2]) const {
        const double s = std::sin(y[1]), c = std::cos(y[1]);
        Gm[0][0] = c;
        Gm[0][1] = -y[0] * s;
        Gm[1][0] = s;
        Gm[1][1] = y[0] * c;
    }
};

template <int N, class P, class C> struct ReparamProb {
    using R = typename C::R;
    P in;
    C ch;
    long nf = 0;
    Vec<R, N> F(const Vec<R, N>& y) {
        ++nf;
        return in.F(ch.g(y));
    }
    Mat<R, N> J(const Vec<R, N>& y) const {
        const Mat<R, N> Ji = in.J(ch.g(y));
        R Gm[N][N];
        ch.G(y, Gm);
        Mat<R, N> Jm;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                R acc = Ji[i] * Gm[0][j];
                for (int k = 1; k < N; ++k) acc = acc + Ji[i + k * N] * Gm[k][j];
                Jm[i + j * N] = acc;
            }
        return Jm;
    }
};

}  // namespace orc
