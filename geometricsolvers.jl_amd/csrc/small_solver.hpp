// The lane-per-system solve: one work-item owns one system of N ≤ 8 unknowns, with the iterate, the
// residual, the Jacobian and its LU factors in VGPRs. It fuses the whole R1 loop of the reference —
// `solve!!` (src/nonlinear/solver.jl:277-378) with Newton's plain and mixed-precision steps
// (src/nonlinear/newton.jl:1302-1485, 1604-1720), the static LU with power-of-two scaling
// (src/linear/static/lu.jl:24-92, src/linear/factors.jl:802-949), iterative refinement
// (src/nonlinear/refinement.jl:211-244) and the Static / Backtracking globalizations
// (src/globalization/linefunction.jl:114-199, linesearch.jl:154-258, backtracking.jl:62-132) — into
// straight-line device code, so that one launch solves the whole batch with no host round trip.
//
// Compile-time choices (template parameters): the family (N, F, J), the globalization, the factor
// precision, refinement, and whether the status diagnostics are produced. Everything else in the
// method and the options is a runtime value.
#pragma once

#include <type_traits>

#include "device_common.hpp"
#include "families.hpp"

namespace gs {

enum : uint8_t { kSuccess = 0, kStalled = 1, kMaxIters = 2, kSingular = 3, kNonfinite = 4, kLsFailed = 5 };

// The runtime part of the method and of Options, in the working precision T (Options{R} and the
// method adapted with ToReal{R}, whose floats are R(x)).
template <class T> struct DevParamsT {
    // Options (src/base/options.jl:58-71)
    T f_abstol, f_reltol, x_abstol, x_reltol, kappa;
    int32_t min_iterations, max_iterations, max_stalls;
    int32_t norm_max;      // 1 = MaxNorm
    int32_t roundoff, rate_test;
    // method
    T alpha;               // Static(α)
    T c1, p;               // Backtracking
    int32_t bt_maxiter;
    int32_t ref_maxiter;
    T ref_eta_eff;         // max(η, √eps)
    T mu;                  // regularization
    int32_t best;
    int32_t reuse;         // Newton(; reuse = k): k ≥ 1 (the REUSE kernels only)
    T dl_radius;           // DogLeg(; radius): the radius each solve starts from, 0 = unset
};
using DevParams = DevParamsT<double>;

enum { kStatic = 0, kBacktracking = 1, kDogLeg = 2 };

// The outcome of one step (StepInfo, src/base/status.jl:220-234).
template <class T> struct StepT {
    uint8_t code;
    bool promoted, shortened;
    int32_t failures;
    T fnorm, stepnorm;
    bool limited;  // a trust-region radius, not the distance to a root, set the step's norm
};

// ------------------------------------------------------------------------------------------------
// Static LU in registers
// ------------------------------------------------------------------------------------------------

// LU with partial pivoting of A (already divided by its power of two), first-largest |a| pivot, row
// exchanges of whole rows, exact multipliers a/p, rank-1 updates without FMA; info = first zero pivot.
// The row exchange is a select per entry; a wave in which no lane exchanges skips it.
template <class TF, int N> GS_DEV int lu_inplace(TF (&A)[N][N], int (&piv)[N]) {
    int info = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int p = k;
        TF amax = tabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            TF a = tabs(A[i][k]);
            if (a > amax) {
                amax = a;
                p = i;
            }
        }
        piv[k] = p;
        if (k + 1 < N && __any(p != k)) {
#pragma unroll
            for (int i = k + 1; i < N; ++i) {
                bool s = (p == i);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    TF t = A[k][j];
                    A[k][j] = s ? A[i][j] : t;
                    A[i][j] = s ? t : A[i][j];
                }
            }
        }
        TF akk = A[k][k];
        if (akk == TF(0)) {
            if (info == 0) info = k + 1;
        } else {
            // a / 1 is a itself (IEEE division by one is exact): a wave whose pivots in this column
            // are all exactly 1 keeps the column as its multipliers
            if (!__all(akk == TF(1))) {
#pragma unroll
                for (int i = k + 1; i < N; ++i) A[i][k] = A[i][k] / akk;
            }
#pragma unroll
            for (int j = k + 1; j < N; ++j)
#pragma unroll
                for (int i = k + 1; i < N; ++i) A[i][j] = A[i][j] - A[i][k] * A[k][j];
        }
    }
    return info;
}

// backsolve (lu.jl:67-92): permutation, unit-lower forward, upper backward substitution.
template <class TF, int N> GS_DEV void lu_backsolve(const TF (&A)[N][N], const int (&piv)[N], TF (&x)[N]) {
#pragma unroll
    for (int k = 0; k < N - 1; ++k) {
        int p = piv[k];
        if (__any(p != k)) {
#pragma unroll
            for (int i = k + 1; i < N; ++i) {
                bool s = (p == i);
                TF t = x[k];
                x[k] = s ? x[i] : t;
                x[i] = s ? t : x[i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        TF xk = x[k];
#pragma unroll
        for (int i = k + 1; i < N; ++i) x[i] = x[i] - A[i][k] * xk;
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        TF xk = x[k];
        if (!__all(A[k][k] == TF(1))) xk = xk / A[k][k];  // x / 1 is x itself
        x[k] = xk;
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = x[i] - A[i][k] * xk;
    }
}

// The factor of a square matrix in TF: J/2^e, pivots, info, e (StaticLU, factors.jl:134-148).
template <class TF, int N> struct Factor {
    TF A[N][N];
    int piv[N];
    int info;
    int e;
};

// factorize!! (lu.jl:24-64) after factor_matrix (factors.jl:802-811): in the working precision only
// e < 0 divides; a narrower TF divides for every e ≠ 0.
template <class TF, int N, class T> GS_DEV void factorize(const T (&J)[N][N], Factor<TF, N>& f) {
    T flat[N * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) flat[i * N + j] = J[i][j];
    constexpr bool same = sizeof(TF) == sizeof(T);
    // A largest finite |entry| of at least 1 (below 2 for a narrower factor) gives e ≥ 0 (e = 0), and
    // any NaN e = 0 as well, so no scaling: a wave whose lanes all have one skips the exponent scan.
    const T m = max_abs<N * N>(flat);
    const bool e0 = m >= T(1) && (same || m < T(2));
    const int e = __all(e0) ? 0 : max_exponent<N * N>(flat, Lim<T>::emax);
    bool scale = e != 0 && (!same || e < 0);
    f.e = scale ? e : 0;
    // the prescale is exact, and a wave in which no lane scales skips it
    const int s = scale ? -e : 0;
    if (__any(scale)) {
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) f.A[i][j] = TF(ldexp2(J[i][j], s));
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) f.A[i][j] = TF(J[i][j]);
    }
    f.info = lu_inplace<TF, N>(f.A, f.piv);
}

// ldiv!! in the working precision (factors.jl:931-936).
template <int N, class T> GS_DEV void ldiv_same(const Factor<T, N>& f, const T (&r)[N], T (&d)[N]) {
    // With f.e = 0 the right-hand side's exponent k is 0 unless its largest finite |entry| reaches
    // 2^(emax − headroom + 1) (a NaN or ∞ entry gives 0 too): a wave whose lanes all have f.e = 0
    // and smaller entries solves unscaled without the exponent scan.
    constexpr T kbig = exp2i<T>(Lim<T>::emax - headroom(N) + 1);
    if (__all(f.e == 0 && max_abs<N>(r) < kbig)) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = r[i];
        lu_backsolve<T, N>(f.A, f.piv, d);
        return;
    }
    const int k = rhs_exponent<N>(r, f.e);
    if (__any(k != 0)) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = ldexp2(r[i], -k);  // ldexp by 0 is exact
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = r[i];
    }
    lu_backsolve<T, N>(f.A, f.piv, d);
    const int u = k - f.e;
    if (__any(u != 0)) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = ldexp2(d[i], u);
    }
}

// ldiv!! with FP32 factors and an FP64 right-hand side (factors.jl:943-949).
template <int N> GS_DEV void ldiv_mixed(const Factor<float, N>& f, const double (&r)[N], double (&d)[N]) {
    double s = __builtin_fabs(r[0]);
#pragma unroll
    for (int i = 1; i < N; ++i) s = jmax(s, __builtin_fabs(r[i]));
    if (s == 0.0) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
        return;
    }
    float t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = float(r[i] / s);
    lu_backsolve<float, N>(f.A, f.piv, t);
    int ex = isfin(s) ? __builtin_amdgcn_frexp_exp(s) : 0;
    double frac = isfin(s) ? __builtin_amdgcn_frexp_mant(s) : s;
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = ldexp2(double(t[i]) * frac, ex - f.e);
}

template <int N, class T> GS_DEV void ldiv(const Factor<T, N>& f, const T (&r)[N], T (&d)[N]) { ldiv_same<N>(f, r, d); }
template <int N> GS_DEV void ldiv(const Factor<float, N>& f, const double (&r)[N], double (&d)[N]) { ldiv_mixed<N>(f, r, d); }

// StaticArrays' SMatrix × SVector: entry i = fma-fold of row i (combine_products' muladd chain).
template <int N, class T> GS_DEV void matvec(const T (&A)[N][N], const T (&v)[N], T (&y)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T acc = A[i][0] * v[0];
#pragma unroll
        for (int j = 1; j < N; ++j) acc = tfma(A[i][j], v[j], acc);
        y[i] = acc;
    }
}

// ‖Jᴴ F‖/(‖J‖_F ‖F‖) (grad_ratio_and_norm, newton.jl:1287-1298).
template <int N, class T> GS_DEV T grad_ratio_scaled(const T (&J)[N][N], const T (&Fx)[N], T jnorm, T fnorm) {
    int e = exponent_of(fnorm);
    T sc = ldexp2(T(1), -e);
    T g[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        T acc = J[0][k] * (Fx[0] * sc);
#pragma unroll
        for (int j = 1; j < N; ++j) acc = tfma(J[j][k], Fx[j] * sc, acc);
        g[k] = acc;
    }
    return rnorm<N>(g) / jnorm / ldexp2(fnorm, -e);
}

template <int N, class T> GS_DEV T grad_ratio(const T (&J)[N][N], const T (&Fx)[N], T fnorm) {
    T flat[N * N];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) flat[i + j * N] = J[i][j];
    T jnorm = rnorm<N * N>(flat);
    if (jnorm == T(0) || fnorm == T(0)) return T(0);
    T g[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        T acc = J[0][k] * Fx[0];
#pragma unroll
        for (int j = 1; j < N; ++j) acc = tfma(J[j][k], Fx[j], acc);
        g[k] = acc;
    }
    T gnorm = rnorm<N>(g);
    if (__builtin_expect(isfin(gnorm) || !(isfin(jnorm) && isfin(fnorm)), 1)) return gnorm / jnorm / fnorm;
    return grad_ratio_scaled<N>(J, Fx, jnorm, fnorm);
}

// backtrack_step (backtracking.jl:320-340): the next trial step from the last two, by quadratic then
// cubic interpolation, kept in [0.1 a, p a].
template <class T> GS_DEV T backtrack_next(T phi0, T d0, T a, T pa, T ap, T pp, T p) {
    if (!isfin(pa)) return T(0.1) * a;
    T an = tnan<T>();
    T s = tabs(d0);
    T r1 = (pa - phi0) / s + a;
    if (!(ap > T(0.0))) {
        if (r1 > T(0.0)) an = (a * a) / (T(2.0) * r1);
    } else if (isfin(pp) && a != ap) {
        T q1 = r1 / (a * a);
        T q2 = ((pp - phi0) / s + ap) / (ap * ap);
        T A = (q1 - q2) / (a - ap);
        T B = (a * q2 - ap * q1) / (a - ap);
        T disc = B * B + T(3.0) * A;
        if (disc >= T(0.0)) an = B > T(0.0) ? T(1.0) / (B + tsqrt(disc)) : (-B + tsqrt(disc)) / (T(3.0) * A);
    }
    if (!(isfin(an) && an > T(0.0))) an = p * a;
    return jmin(jmax(an, T(0.1) * a), p * a);
}

// ------------------------------------------------------------------------------------------------
// One system
// ------------------------------------------------------------------------------------------------

// EAGER: the status's gradient ratio is formed at every factorisation (cheaper where a solve usually
// ends after its first step); otherwise once, at exit, at the last factorisation's point.
template <class Fam, int GLOB, bool MIXED, bool REFINE, bool DIAG, bool REUSE = false, bool EAGER = false>
struct SmallSolve {
    static constexpr int N = Fam::N;
    using T = typename Fam::T;  // the working precision
    using TF = typename std::conditional<MIXED, float, T>::type;
    using Step = StepT<T>;
    const Fam& fam;
    const DevParamsT<T>& q;

    // the per-solve state (NewtonState, newton.jl:715-744, as the R1 value state holds it)
    T merit;       // norm2(Fx) for Backtracking
    T sfnorm;      // ‖Fx‖ that the state carries
    T gradratio;   // diagnostics of the last factorisation
    // The point of the last factorisation's Jacobian, the residual there and the ‖F‖ its diagnostics
    // read: only the last factorisation's gradient ratio reaches the status, so it is formed once, at
    // exit, from that residual and J evaluated again at that point (J is a function of x alone, so
    // the values are the same).
    T jx[N], jF[N];  // jF: F at jx, the residual the factorisation read
    T jfn;
    bool have_j;
    bool promoted;      // mixed precision: factorising in FP64
    int slow_steps;
    // reuse = k: the factor the state holds, in the factor precision, and the rule state FactorAge
    // (newton.jl:621-633): its age, whether the last step stalled, whether it stepped along a reused
    // factor
    Factor<TF, N> kept;
    int age;
    bool stalled, reused;
    // DogLeg: the trust radius (TrustRadius, newton.jl:637-640), 0 until the first step sets it
    T radius;

    GS_DEV SmallSolve(const Fam& f, const DevParamsT<T>& p) : fam(f), q(p) {}

    // residual_norm (newton.jl:1274-1278)
    GS_DEV T residual_norm(const T (&Fx)[N]) const {
        if (GLOB == kBacktracking) {
            if (isfin(merit) && merit >= T(N) * Lim<T>::min) return tsqrt(merit);
            return rnorm<N>(Fx);
        }
        return isnan_(sfnorm) ? rnorm<N>(Fx) : sfnorm;
    }

    // --- globalization --------------------------------------------------------------------------

    // Static (linefunction.jl:114-131).
    GS_DEV Step step_static(T (&x)[N], T (&Fx)[N], const T (&d)[N]) const {
        T s = q.alpha;
        T dn = rnorm<N>(d);
        T stepnorm = tabs(s) * dn;
        if (!isfin(stepnorm)) return Step{kNonfinite, false, false, 0, rnorm<N>(Fx), stepnorm};
        T xp[N], Fp[N];
#pragma unroll
        for (int i = 0; i < N; ++i) xp[i] = x[i] + s * d[i];
        fam.F(xp, Fp);
        T fn = rnorm<N>(Fp);
        if (!isfin(fn)) return Step{kNonfinite, false, false, 0, rnorm<N>(Fx), stepnorm};
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = xp[i];
            Fx[i] = Fp[i];
        }
        return Step{kSuccess, false, false, 0, fn, stepnorm};
    }

    GS_DEV static T backtrack_step(T phi0, T d0, T a, T pa, T ap, T pp, T p) {
        return backtrack_next<T>(phi0, d0, a, pa, ap, pp, p);
    }

    // globalize_merit!! for Backtracking (linefunction.jl:159-199 → linesearch.jl:154-171 →
    // backtracking.jl:354-390). `exact`: ExactStep (slope −2φ₀) or MeasuredSlope (φ′(0) = 2⟨F, J d⟩).
    GS_DEV Step step_backtracking(T (&x)[N], T (&Fx)[N], const T (&d)[N], bool exact) {
        const T phi0 = merit;
        T dn = rnorm<N>(d);
        if (!isfin(dn)) return Step{kNonfinite, false, false, 0, rnorm<N>(Fx), dn};
        T d0;
        if (exact) {
            d0 = -T(2.0) * phi0;
        } else {
            T xa[N], Fa[N], Jv[N];
#pragma unroll
            for (int i = 0; i < N; ++i) xa[i] = x[i] + T(0.0) * d[i];
            fam.F(xa, Fa);
            if constexpr (is_deflated<Fam>::value) {
                fam.Jv(xa, d, Jv);  // the deflated product M (J d) + ⟨∇M, d⟩ F (deflation.jl:352-357)
            } else {
                T Ja[N][N];
                fam.J(xa, Ja);
                matvec<N>(Ja, d, Jv);
            }
            d0 = T(2.0) * rdot<N>(Fa, Jv);
        }
        if (!(isfin(phi0) && isfin(d0) && d0 < T(0.0))) {
            uint8_t c = !(isfin(phi0) && isfin(d0)) ? kNonfinite : (d0 > T(0.0) ? kLsFailed : kStalled);
            // a LINESEARCH_FAILED anchor keeps x with φ = NaN: the failure takes no step
            if (c == kLsFailed) return Step{kLsFailed, false, false, 1, rnorm<N>(Fx), T(0.0)};
            return Step{c, false, false, 0, rnorm<N>(Fx), T(0.0)};
        }
        const T tau = tfmax(T(4.0) * Lim<T>::eps * tabs(phi0), T(4.0) * Lim<T>::tiny);
        const T amin = jmax(tau / tabs(d0), Lim<T>::min);
        T alpha = tfmin(tfmax(T(1.0), amin), T(__builtin_inf()));
        T aa = alpha, pa = tnan<T>();
        T Fa[N];
        T ap = tnan<T>(), pp = tnan<T>();
        int frozen = 0;
        uint8_t code = kLsFailed;
        const T sqrteps = Lim<T>::sqrteps;
        for (int trips = 0; trips < q.bt_maxiter; ++trips) {
            aa = alpha;
            T xa[N];
#pragma unroll
            for (int i = 0; i < N; ++i) xa[i] = x[i] + alpha * d[i];
            fam.F(xa, Fa);
            pa = norm2<N>(Fa);
            frozen = (pa == phi0 && alpha <= sqrteps) ? frozen + 1 : 0;
            if (pa - phi0 <= jmin(T(0.0), q.c1 * alpha * d0 + tau)) {
                code = (isfin(pa) && pa - phi0 < -tau) ? kSuccess : (tabs(pa - phi0) <= tau ? kStalled : kLsFailed);
                break;
            } else if (frozen >= 2) {
                code = kStalled;
                break;
            } else if (alpha <= amin) {
                code = tabs(pa - phi0) <= tau ? kStalled : kLsFailed;
                break;
            }
            T an = isfin(pa) ? backtrack_step(phi0, d0, alpha, pa, ap, pp, q.p)
                                  : jmin(T(0.1) * alpha, tsqrt(alpha) * tsqrt(amin));
            ap = alpha;
            pp = pa;
            alpha = jmax(an, amin);
        }
        if (code == kSuccess) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                x[i] = x[i] + aa * d[i];
                Fx[i] = Fa[i];
            }
            merit = pa;
            return Step{kSuccess, false, aa < T(1.0), 0, tsqrt(pa), aa * dn};
        }
        if (code == kLsFailed) {
            if (isfin(pa) && pa < phi0) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    x[i] = x[i] + aa * d[i];
                    Fx[i] = Fa[i];
                }
                merit = pa;
                return Step{kLsFailed, false, aa < T(1.0), 1, tsqrt(pa), aa * dn};
            }
            return Step{kLsFailed, false, false, 1, rnorm<N>(Fx), T(0.0)};
        }
        return Step{code == kStalled ? kStalled : kNonfinite, false, false, 0, rnorm<N>(Fx), T(0.0)};
    }

    GS_DEV Step globalize(T (&x)[N], T (&Fx)[N], const T (&d)[N], bool exact) {
        if (GLOB == kStatic) return step_static(x, Fx, d);
        return step_backtracking(x, Fx, d, exact);
    }

    // --- refinement -----------------------------------------------------------------------------

    // The product of a refinement (newton.jl:1344-1384): the held regularised Jacobian J of a fresh
    // factor, or, along a reused factor (JVP), the JVP at x with J = J(x): 1·J(x) v + μ v where μ > 0
    // (RegularizedJVP). A Deflated family's JVP is its own product M (J v) + ⟨∇M, v⟩ F at x
    // (deflation.jl:352-357), not the deflated Jacobian times v.
    template <bool JVP> GS_DEV void product(const T (&J)[N][N], const T (&x)[N], const T (&v)[N], T (&Jv)[N]) const {
        if constexpr (JVP && is_deflated<Fam>::value) {
            fam.Jv(x, v, Jv);
        } else {
            matvec<N>(J, v, Jv);
        }
        if (JVP && q.mu != T(0.0)) {
#pragma unroll
            for (int i = 0; i < N; ++i) Jv[i] = T(1.0) * Jv[i] + q.mu * v[i];
        }
    }

    // refine_passes!! (refinement.jl:211-244) with the product of `product<JVP>(J, ·)`.
    // `rlast` receives ‖r‖ = ‖−F − J d‖ of the returned d when the refinement succeeds; x is the
    // point of the JVP (along a reused factor).
    template <class TFR, bool JVP = false>
    GS_DEV void refine(T (&d)[N], const Factor<TFR, N>& f, const T (&J)[N][N], const T (&Fx)[N], const T (&x)[N],
                       bool& failed) const {
        T rlast;
        refine<TFR, JVP>(d, f, J, Fx, x, failed, rlast);
    }
    template <class TFR, bool JVP = false>
    GS_DEV void refine(T (&d)[N], const Factor<TFR, N>& f, const T (&J)[N][N], const T (&Fx)[N], const T (&x)[N],
                       bool& failed, T& rlast) const {
        rlast = tnan<T>();
        failed = false;
        // Under Static without reuse the state's ‖F‖ is rnorm of this very residual (set from it at the
        // start and by every step), so it is taken instead of folding F again.
        T fnorm = (GLOB == kStatic && !REUSE) ? residual_norm(Fx) : rnorm<N>(Fx);
        if (fnorm == T(0.0)) return;
        if (!isfin(fnorm)) {
            failed = true;
            return;
        }
        T Jv[N], r[N];
        product<JVP>(J, x, d, Jv);
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] = -Fx[i] + -Jv[i];
        T rprev = rnorm<N>(r);
        if (!isfin(rprev)) {
            failed = true;
            return;
        }
        rlast = rprev;
        if (rprev <= q.ref_eta_eff * fnorm) return;
        for (int it = 1; it <= q.ref_maxiter; ++it) {
            T dd[N], dc[N];
            ldiv<N>(f, r, dd);
#pragma unroll
            for (int i = 0; i < N; ++i) dc[i] = d[i] + dd[i];
            product<JVP>(J, x, dc, Jv);
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = -Fx[i] + -Jv[i];
            T rcur = rnorm<N>(r);
            if (!(rcur < rprev)) {
                failed = true;
                return;
            }
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = dc[i];
            rprev = rcur;
            rlast = rcur;
            if (rcur <= q.ref_eta_eff * fnorm) return;
        }
        failed = true;
    }

    // --- steps ----------------------------------------------------------------------------------

    GS_DEV void jacobian(const T (&x)[N], const T (&Fx)[N], T (&J)[N][N]) {
        if (DIAG && EAGER) {
            const T fnorm = residual_norm(Fx);
            fam.J(x, J);
            gradratio = grad_ratio<N>(J, Fx, fnorm);
        } else {
            if (DIAG) {
                jfn = residual_norm(Fx);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    jx[i] = x[i];
                    jF[i] = Fx[i];
                }
                have_j = true;
            }
            fam.J(x, J);
        }
        if (q.mu != T(0.0)) {
#pragma unroll
            for (int i = 0; i < N; ++i) J[i][i] = J[i][i] + q.mu;
        }
    }

    // newton_step!! in the working precision (newton.jl:1679-1688, 1302-1311, 1396-1416).
    GS_DEV Step plain_step(T (&x)[N], T (&Fx)[N]) {
        T J[N][N];
        jacobian(x, Fx, J);
        Factor<T, N> f;
        factorize<T, N>(J, f);
        Step s;
        if (f.info > 0) {
            s = Step{kSingular, false, false, 0, rnorm<N>(Fx), T(0.0)};
        } else {
            T d[N];
            ldiv<N>(f, Fx, d);
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            if (REFINE) {
                bool failed;
                refine<T>(d, f, J, Fx, x, failed);
            }
            bool exact = q.mu == T(0.0) && !REFINE;
            s = globalize(x, Fx, d, exact);
        }
        sfnorm = s.fnorm;
        return s;
    }

    // mixed_newton!! with a fresh factor and no store (newton.jl:1604-1666): the FP32 factor and its
    // refinement, or the FP64 one of a promoted state; a failed FP32 factor or refinement promotes
    // and refactorises the held Jacobian in FP64. J receives the regularised Jacobian; returns
    // whether a direction exists (md.ok).
    GS_DEV bool mixed_newton(const T (&x)[N], const T (&Fx)[N], T fnorm_old, T (&J)[N][N],
                             T (&d)[N], int& failures, T& lin_rel) {
        jacobian(x, Fx, J);
        bool ok = true;
        failures = 0;
        lin_rel = T(0.0);
        if (promoted) {
            Factor<T, N> f;
            factorize<T, N>(J, f);
            if (f.info > 0) {
                ok = false;
            } else {
                ldiv<N>(f, Fx, d);
#pragma unroll
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                if (REFINE) {
                    bool failed;
                    refine<T>(d, f, J, Fx, x, failed);
                }
            }
        } else {
            Factor<float, N> f;
            factorize<float, N>(J, f);
            bool ref_failed = false;
            T rlast = tnan<T>();
            if (f.info > 0) {
                ref_failed = true;
            } else {
                ldiv<N>(f, Fx, d);
#pragma unroll
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                if (REFINE) refine<float>(d, f, J, Fx, x, ref_failed, rlast);
            }
            if (ref_failed) {
                // promote_factor!! (newton.jl:1332-1339): refactorise the held Jacobian in FP64
                failures += 1;
                promoted = true;
                Factor<T, N> g;
                factorize<T, N>(J, g);
                if (g.info > 0) {
                    ok = false;
                } else {
                    ldiv<N>(g, Fx, d);
#pragma unroll
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                }
            } else if (fnorm_old != T(0.0) && isfin(fnorm_old)) {
                // linear_residual_ratio (newton.jl:1422-1430): ‖F + J d‖ / ‖F‖. A successful refinement
                // has just formed r = −F − J d of this d and its norm, which has the same bits (−a + −b
                // = −(a + b) exactly, and the norm of −v is that of v).
                T lin_norm;
                if (REFINE) {
                    lin_norm = rlast;
                } else {
                    T Jd[N], lin[N];
                    matvec<N>(J, d, Jd);
#pragma unroll
                    for (int i = 0; i < N; ++i) lin[i] = T(1.0) * Fx[i] + T(1.0) * Jd[i];
                    lin_norm = rnorm<N>(lin);
                }
                lin_rel = lin_norm / fnorm_old;
            }
        }
        return ok;
    }

    // read_mixed_stall (newton.jl:1437-1451), η_lin = max(√eps, T(1e-4)): two slow steps whose FP32
    // linear model is inaccurate promote.
    GS_DEV void read_mixed_stall(Step& s, T fnorm_old, T lin_rel) {
        if (s.fnorm > T(0.5) * fnorm_old && lin_rel > T(1e-4)) {
            int slow = slow_steps + 1;
            if (slow >= 2) {
                promoted = true;
                slow_steps = 0;
                s.failures += 1;
                s.promoted = true;
            } else {
                slow_steps = slow;
            }
        } else {
            slow_steps = 0;
        }
    }

    // mixed_step!! / mixed_newton!! with fresh factors (newton.jl:1604-1666, 1692-1720).
    GS_DEV Step mixed_step(T (&x)[N], T (&Fx)[N]) {
        T fnorm_old = residual_norm(Fx);
        const bool was_promoted = promoted;
        T J[N][N], d[N];
        int failures;
        T lin_rel;
        if (!mixed_newton(x, Fx, fnorm_old, J, d, failures, lin_rel)) {
            Step s{kSingular, promoted, false, failures, rnorm<N>(Fx), T(0.0)};
            return s;
        }
        const bool promoted_step = promoted && !was_promoted;
        const bool exact = q.mu == T(0.0) && promoted;
        Step s = globalize(x, Fx, d, exact);
        s.failures += failures;
        if (promoted_step) {
            s.promoted = true;
            slow_steps = 0;
        } else if (!was_promoted) {
            read_mixed_stall(s, fnorm_old, lin_rel);
        }
        sfnorm = s.fnorm;
        return s;
    }

    // --- DogLeg (dogleg.jl:116-206, newton.jl:1725-1780) ----------------------------------------

    GS_DEV static T dl_clamp(T v) { return v > T(1e10) ? T(1e10) : (v < T(1e-6) ? T(1e-6) : v); }
    GS_DEV static T dl_shrink(T v) { return jmax(T(0.25) * v, T(1e-6)); }
    GS_DEV static T dl_grow(T v) { return jmin(T(4.0) * v, T(1e10)); }

    // The trial loop of dogleg_step on the model matrix J (the regularised Jacobian that was
    // factorised) with the Newton step sN (read where `newton`): the Newton step inside the radius,
    // else the Cauchy point cut to the radius, else the point of norm δ on the segment between them;
    // a trial is accepted when pred > 0 and ared ≥ T(1e-4) pred (unsquared norms), a rejection shrinks δ
    // to max(δ/4, T(1e-6)), and a rejection at T(1e-6) fails. x, Fx and `radius` are updated.
    GS_DEV Step dogleg_trial(T (&x)[N], T (&Fx)[N], const T (&J)[N][N], const T (&sN)[N],
                             bool newton) {
        const T lo = T(1e-6);
        const T fnorm = rnorm<N>(Fx);
        if (fnorm == T(0.0)) return Step{kSuccess, false, false, 0, fnorm, T(0.0)};
        const T nN = newton ? rnorm<N>(sN) : T(__builtin_inf());
        // the Cauchy direction from the unit residual: g̃ = Jᵀ (F/‖F‖), ĝ = g̃/‖g̃‖
        T gh[N];
        const T finv = T(1.0) / fnorm;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            T acc = J[0][k] * (finv * Fx[0]);
#pragma unroll
            for (int j = 1; j < N; ++j) acc = tfma(J[j][k], finv * Fx[j], acc);
            gh[k] = acc;
        }
        const T gnorm = rnorm<N>(gh);
        const T ginv = T(1.0) / gnorm;
#pragma unroll
        for (int i = 0; i < N; ++i) gh[i] = ginv * gh[i];
        T Jg[N];
        matvec<N>(J, gh, Jg);
        const T jnorm = rnorm<N>(Jg);
        const bool cauchy = gnorm != T(0.0) && isfin(gnorm) && jnorm != T(0.0) && isfin(jnorm);
        const T nC = fnorm * ((gnorm / jnorm) / jnorm);
        T delta = radius;
        if (delta == T(0.0) && (newton || cauchy)) {
            const T n0 = newton ? nN : nC;
            delta = n0 < lo ? T(2.0) * lo : dl_clamp(n0);
        }
        while (true) {
            T sv[N];
            bool boundary;
            if (newton && nN <= delta) {
#pragma unroll
                for (int i = 0; i < N; ++i) sv[i] = sN[i];
                boundary = nN >= delta;
            } else if (!cauchy) {
                radius = delta;
                return Step{kStalled, false, false, 0, fnorm, T(0.0)};
            } else if (!newton || nC >= delta) {
                const T a = -jmin(nC, delta);
#pragma unroll
                for (int i = 0; i < N; ++i) sv[i] = a * gh[i];
                boundary = nC >= delta;
            } else {
                // the root t ≥ 0 of ‖a + t b̂‖ = 1 in units of δ, a = −(nC/δ) ĝ, b = sN + nC ĝ
                T av[N], bh[N];
                const T ac = -(nC / delta);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    av[i] = ac * gh[i];
                    bh[i] = T(1.0) * sN[i] + nC * gh[i];
                }
                const T binv = T(1.0) / rnorm<N>(bh);
#pragma unroll
                for (int i = 0; i < N; ++i) bh[i] = binv * bh[i];
                const T ab = rdot<N>(av, bh);
                const T c = T(1.0) - norm2<N>(av);
                const T ab2 = ab * ab;
                const T t = c / (ab + tsqrt(ab2 + c));
                const T dt = delta * t;
#pragma unroll
                for (int i = 0; i < N; ++i) sv[i] = delta * av[i] + dt * bh[i];
                boundary = true;
            }
            T Js[N], Flin[N], xp[N], Fp[N];
            matvec<N>(J, sv, Js);
#pragma unroll
            for (int i = 0; i < N; ++i) Flin[i] = T(1.0) * Fx[i] + T(1.0) * Js[i];
            const T pred = fnorm - rnorm<N>(Flin);
#pragma unroll
            for (int i = 0; i < N; ++i) xp[i] = x[i] + T(1.0) * sv[i];
            fam.F(xp, Fp);
            const T fnp = rnorm<N>(Fp);
            const T ared = fnorm - fnp;
            if (pred > T(0.0) && ared >= T(1e-4) * pred) {
                const bool limited = !(newton && nN <= delta);
                const T rho = ared / pred;
                if (rho < T(0.1)) {
                    delta = nN < delta ? jmax(nN, lo) : dl_shrink(delta);
                } else if (rho > T(0.75) && boundary) {
                    delta = dl_grow(delta);
                }
                const T stepnorm = rnorm<N>(sv);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    x[i] = xp[i];
                    Fx[i] = Fp[i];
                }
                radius = delta;
                return Step{kSuccess, false, false, 0, fnp, stepnorm, limited};
            }
            if (delta == lo) {
                radius = delta;
                return Step{kLsFailed, false, false, 1, fnorm, T(0.0)};
            }
            delta = dl_shrink(delta);
        }
    }

    // newton_step!! with DogLeg (newton.jl:1725-1737) and dogleg_direction (newton.jl:1741-1746): a
    // failed factor, or a Newton step of zero or non-finite norm, leaves the Cauchy step.
    GS_DEV Step plain_step_dogleg(T (&x)[N], T (&Fx)[N]) {
        T J[N][N];
        jacobian(x, Fx, J);
        Factor<T, N> f;
        factorize<T, N>(J, f);
        T d[N];
        bool newton = false;
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = Fx[i];
        if (f.info <= 0) {
            ldiv<N>(f, Fx, d);
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            if (REFINE) {
                bool failed;
                refine<T>(d, f, J, Fx, x, failed);
            }
            const T nd = rnorm<N>(d);
            newton = isfin(nd) && nd != T(0.0);
        }
        Step s = dogleg_trial(x, Fx, J, d, newton);
        sfnorm = s.fnorm;
        return s;
    }

    // mixed_step!! with DogLeg (newton.jl:1757-1780).
    GS_DEV Step mixed_step_dogleg(T (&x)[N], T (&Fx)[N]) {
        const T fnorm_old = residual_norm(Fx);
        const bool was_promoted = promoted;
        T J[N][N], d[N];
        int failures;
        T lin_rel;
        const bool ok = mixed_newton(x, Fx, fnorm_old, J, d, failures, lin_rel);
        const bool promoted_step = promoted && !was_promoted;
        bool newton = false;
        if (ok) {
            const T nd = rnorm<N>(d);
            newton = isfin(nd) && nd != T(0.0);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = Fx[i];
        }
        Step s = dogleg_trial(x, Fx, J, d, newton);
        s.failures += failures;
        if (promoted_step) {
            s.promoted = true;
            slow_steps = 0;
        } else if (!was_promoted) {
            read_mixed_stall(s, fnorm_old, lin_rel);
        }
        sfnorm = s.fnorm;
        return s;
    }

    // --- reuse = k (newton.jl:1781-1874) ---------------------------------------------------------

    GS_DEV int fresh_age() const { return q.reuse == 1 ? 0 : 1; }                // newton.jl:629
    GS_DEV int next_age() const { return age + 1 >= q.reuse ? 0 : age + 1; }      // newton.jl:632

    // reuse_result (newton.jl:1814-1820): a SINGULAR step resets the age, a STALLED one refreshes
    // the factor at the next step.
    GS_DEV void set_rule(uint8_t code, int age_next, bool along_reused) {
        age = code == kSingular ? 0 : age_next;
        stalled = code != kSingular && code == kStalled;
        reused = along_reused;
    }

    // newton_step!! with reuse in the working precision (newton.jl:1781-1810): a step along the held
    // factor refines with the JVP at x and takes the measured slope; a failed refinement (one
    // step-rule failure) or a line search that fails without a step refactorises at x in the step.
    GS_DEV Step plain_step_reuse(T (&x)[N], T (&Fx)[N]) {
        int failures = 0;
        if (!(age == 0 || stalled)) {
            const int age_next = next_age();
            T d[N];
            ldiv<N>(kept, Fx, d);
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            bool ref_failed = false;
            if (REFINE) {
                T Jx[N][N];
                fam.J(x, Jx);
                refine<TF, true>(d, kept, Jx, Fx, x, ref_failed);
            }
            if (ref_failed) {
                failures = 1;
            } else {
                // a failed search without a step leaves x, Fx and the merit as they were
                Step s = globalize(x, Fx, d, false);
                if (!(s.code == kLsFailed && s.stepnorm == T(0.0))) {
                    sfnorm = s.fnorm;
                    set_rule(s.code, age_next, true);
                    return s;
                }
            }
        }
        T J[N][N];
        jacobian(x, Fx, J);
        factorize<TF, N>(J, kept);
        Step s;
        if (kept.info > 0) {
            s = Step{kSingular, false, false, 0, rnorm<N>(Fx), T(0.0)};
        } else {
            T d[N];
            ldiv<N>(kept, Fx, d);
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            if (REFINE) {
                bool failed;
                refine<TF>(d, kept, J, Fx, x, failed);
            }
            s = globalize(x, Fx, d, q.mu == T(0.0) && !REFINE);
        }
        s.failures += failures;
        sfnorm = s.fnorm;
        set_rule(s.code, fresh_age(), false);
        return s;
    }

    // mixed_step!! with reuse (newton.jl:1823-1874) over mixed_newton!! with store = true
    // (newton.jl:1604-1666): a promoted state factorises in FP64 at every step; otherwise the step
    // runs along the held FP32 factor until its age or a stall refreshes it, and a reused direction
    // whose refinement fails is refactorised in FP32 at x (one failure) before any promotion. A line
    // search along the held factor that fails without a step is retried once from a fresh factor.
    GS_DEV Step mixed_step_reuse(T (&x)[N], T (&Fx)[N]) {
        const T fnorm_old = residual_norm(Fx);
        const bool state_promoted = promoted;
        bool fresh = promoted || age == 0 || stalled;
        const int age_fresh = fresh_age();
        int age_next = fresh ? age_fresh : next_age();
        while (true) {
            const bool was_fresh = fresh;
            T d[N];
            bool ok = true, md_fresh = true, have = false;
            int failures = 0;
            T lin_rel = T(0.0);
            if (promoted) {
                T J[N][N];
                jacobian(x, Fx, J);
                Factor<T, N> f;
                factorize<T, N>(J, f);
                if (f.info > 0) {
                    ok = false;
                } else {
                    ldiv<N>(f, Fx, d);
#pragma unroll
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                    if (REFINE) {
                        bool failed;
                        refine<T>(d, f, J, Fx, x, failed);
                    }
                }
                have = true;
            } else if (!fresh) {
                ldiv<N>(kept, Fx, d);
#pragma unroll
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                bool ref_failed = false;
                if (REFINE) {
                    T Jx[N][N];
                    fam.J(x, Jx);
                    refine<TF, true>(d, kept, Jx, Fx, x, ref_failed);
                }
                if (!ref_failed) {
                    md_fresh = false;
                    have = true;
                }
            }
            if (!have) {
                T J[N][N];
                jacobian(x, Fx, J);
                Factor<float, N> f;
                factorize<float, N>(J, f);
                if (!fresh) failures += 1;
                bool ref_failed = false;
                if (f.info > 0) {
                    ref_failed = true;
                } else {
                    ldiv<N>(f, Fx, d);
#pragma unroll
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                    if (REFINE) refine<float>(d, f, J, Fx, x, ref_failed);
                }
                if (ref_failed) {
                    // promote_factor!! (newton.jl:1332-1339): refactorise the held Jacobian in FP64
                    failures += 1;
                    promoted = true;
                    Factor<T, N> g;
                    factorize<T, N>(J, g);
                    if (g.info > 0) {
                        ok = false;
                    } else {
                        ldiv<N>(g, Fx, d);
#pragma unroll
                        for (int i = 0; i < N; ++i) d[i] = -d[i];
                    }
                } else {
                    copy_factor(f);
                    if (fnorm_old != T(0.0) && isfin(fnorm_old)) {
                        // linear_residual_ratio (newton.jl:1422-1430)
                        T Jd[N], lin[N];
                        matvec<N>(J, d, Jd);
#pragma unroll
                        for (int i = 0; i < N; ++i) lin[i] = T(1.0) * Fx[i] + T(1.0) * Jd[i];
                        lin_rel = rnorm<N>(lin) / fnorm_old;
                    }
                }
            }
            fresh = md_fresh;
            if (!ok) {
                set_rule(kSingular, 0, !md_fresh);
                return Step{kSingular, promoted, false, failures, rnorm<N>(Fx), T(0.0)};
            }
            const bool promoted_step = promoted && !state_promoted;
            if (promoted_step) {
                age_next = 0;
            } else if (md_fresh && !was_fresh) {
                age_next = age_fresh;
            }
            // newton_stepkind(prob, m, T, md.fresh, promoted) (newton.jl:1244-1254)
            Step s = globalize(x, Fx, d, md_fresh && q.mu == T(0.0) && promoted);
            if (!md_fresh && s.code == kLsFailed && s.stepnorm == T(0.0)) {
                fresh = true;
                age_next = age_fresh;
                continue;
            }
            s.failures += failures;
            if (promoted_step) {
                s.promoted = true;
                slow_steps = 0;
            } else if (md_fresh && !promoted) {
                // read_mixed_stall (newton.jl:1437-1451), η_lin = max(√eps, T(1e-4))
                if (s.fnorm > T(0.5) * fnorm_old && lin_rel > T(1e-4)) {
                    int slow = slow_steps + 1;
                    if (slow >= 2) {
                        promoted = true;
                        slow_steps = 0;
                        s.failures += 1;
                        s.promoted = true;
                    } else {
                        slow_steps = slow;
                    }
                } else {
                    slow_steps = 0;
                }
            }
            sfnorm = s.fnorm;
            set_rule(s.code, age_next, !md_fresh);
            return s;
        }
    }

    // The held factor of a mixed solve is the FP32 one (`kept` is Factor<float, N> when MIXED).
    GS_DEV void copy_factor(const Factor<float, N>& f) {
        if constexpr (MIXED) kept = f;
    }

    GS_DEV Step step(T (&x)[N], T (&Fx)[N]) {
        if constexpr (GLOB == kDogLeg) {
            if constexpr (MIXED) return mixed_step_dogleg(x, Fx);
            else return plain_step_dogleg(x, Fx);
        } else if constexpr (REUSE) {
            if constexpr (MIXED) return mixed_step_reuse(x, Fx);
            else return plain_step_reuse(x, Fx);
        } else {
            if constexpr (MIXED) return mixed_step(x, Fx);
            else return plain_step(x, Fx);
        }
    }

    // --- the loop -------------------------------------------------------------------------------

    struct Result {
        uint8_t code;
        bool contracted, promoted, reused;
        int32_t iterations, failures;
        T fnorm, stepnorm, rate, gradratio;
    };

    // solve!! (solver.jl:277-378) for a square NonlinearProblem with a Newton method.
    // converged_status (solver.jl:122-125): the status fnorm of a step's end at x. A Deflated family
    // reports its inner ‖F‖ = ‖M F‖ / M where the norm is finite (_inner_status, deflation.jl:514-520).
    GS_DEV T status_fnorm(const T (&x)[N], T fn) const {
        if constexpr (is_deflated<Fam>::value) return isfin(fn) ? fn / fam.scale(x) : fn;
        else return fn;
    }
    // test_status (solver.jl:138-146): the norm the residual test reads, from the step's raw fnorm,
    // its status fnorm and the residual Fx at x. Under MaxNorm a Deflated family scales the status
    // fnorm by ‖M F‖∞ / ‖M F‖₂ (deflation.jl:540-548).
    GS_DEV T test_fnorm(const T (&Fx)[N], T raw, T st) const {
        if (!q.norm_max) return st;
        if constexpr (is_deflated<Fam>::value) {
            const T mx = rmaxnorm<N>(Fx);
            return raw == T(0.0) ? mx : (mx / raw) * st;
        } else {
            return rmaxnorm<N>(Fx);
        }
    }

    GS_DEV Result solve(T (&x)[N]) {
        T Fx[N];
        fam.F(x, Fx);
        const T fnorm0 = rnorm<N>(Fx);
        // newton_start!! (newton.jl:1135-1144)
        merit = GLOB == kBacktracking ? norm2<N>(Fx) : tnan<T>();
        sfnorm = fnorm0;
        gradratio = tnan<T>();
        have_j = false;
        promoted = false;
        slow_steps = 0;
        age = 0;
        stalled = false;
        reused = false;
        radius = q.dl_radius;  // reset_rule (newton.jl:677)
        T best_x[N];
        T best_fn = fnorm0;
        if (q.best) {
#pragma unroll
            for (int i = 0; i < N; ++i) best_x[i] = x[i];
        }
        Result R{kSuccess, false, false, false, 0, 0, fnorm0, T(0.0), tnan<T>(), tnan<T>()};
        if (!isfin(fnorm0)) {
            R.code = kNonfinite;
        } else {
            R.fnorm = status_fnorm(x, fnorm0);
            const T ftest0 = test_fnorm(Fx, fnorm0, R.fnorm);
            const T ftol = jmax(q.f_abstol, q.f_reltol * ftest0);
            if (0 >= q.min_iterations && ftest0 <= ftol) {
                R.code = kSuccess;
            } else if (q.max_iterations == 0) {
                R.code = kMaxIters;
            } else {
                int stalls = 0;
                while (true) {
                    const T rate_k = R.rate;
                    Step s = step(x, Fx);
                    if (q.best && s.fnorm < best_fn) {
                        best_fn = s.fnorm;
#pragma unroll
                        for (int i = 0; i < N; ++i) best_x[i] = x[i];
                    }
                    // record (status.jl:245-262)
                    R.rate = R.stepnorm > T(0.0) ? s.stepnorm / R.stepnorm : tnan<T>();
                    R.code = s.code;
                    R.iterations += 1;
                    R.fnorm = status_fnorm(x, s.fnorm);
                    R.stepnorm = s.stepnorm;
                    R.failures += s.failures;
                    R.promoted |= s.promoted;
                    // observe_rate!! (newton.jl:1205-1210): a step of rate θ ≥ 1 along a reused
                    // factor refreshes the factor at the next step
                    if (REUSE && reused && R.rate >= T(1.0)) stalled = true;
                    if (s.code == kSingular || s.code == kNonfinite || s.code == kMaxIters) break;
                    if (s.code == kLsFailed && s.stepnorm == T(0.0)) break;
                    const T ftest = test_fnorm(Fx, s.fnorm, R.fnorm);
                    if (R.iterations >= q.min_iterations && ftest <= ftol) {
                        R.code = kSuccess;
                        break;
                    }
                    // a step that a trust radius limited passes only the residual test
                    const bool steptested = s.code != kStalled && !s.limited;
                    const T xnorm = steptested ? rnorm<N>(x) : tnan<T>();
                    const T xtol = jmax(q.x_abstol, q.x_reltol * xnorm);
                    const bool past_min = R.iterations >= q.min_iterations;
                    bool contracted = false;
                    if (!s.shortened && q.rate_test && past_min) {
                        T th = jmax(R.rate, rate_k);
                        contracted = R.stepnorm > T(0.0) && th < T(1.0) && th / (T(1.0) - th) * R.stepnorm <= q.kappa * xtol;
                    }
                    bool step_conv = past_min && R.stepnorm > T(0.0) && R.stepnorm <= xtol;
                    bool zero_step = s.code == kSuccess && s.stepnorm == T(0.0) && past_min;
                    bool roundoff = q.roundoff && past_min && R.rate >= T(1.0) && rate_k < T(1.0);
                    if (contracted || (steptested && (step_conv || zero_step || roundoff))) {
                        R.code = contracted ? kSuccess : kStalled;
                        R.contracted = contracted;
                        break;
                    }
                    if (s.code == kStalled) {
                        stalls += 1;
                        if (stalls >= q.max_stalls) {
                            R.code = kStalled;
                            break;
                        }
                    } else {
                        stalls = 0;
                    }
                    if (R.iterations >= q.max_iterations) {
                        R.code = kMaxIters;
                        break;
                    }
                }
            }
        }
        if (DIAG && !EAGER && have_j) {
            // grad_ratio_and_norm (newton.jl:1287-1298) of the last factorisation: the residual it
            // read is kept (jF), the Jacobian evaluated again at its point
            T Jj[N][N];
            fam.J(jx, Jj);
            gradratio = grad_ratio<N>(Jj, jF, jfn);
        }
        R.gradratio = gradratio;
        R.reused = REUSE && reused && R.iterations > 0;  // reused_factor (newton.jl:1188-1191)
        // restore_best (newton.jl:792-799)
        if (q.best && R.code != kSuccess && isfin(best_fn) && best_fn < R.fnorm) {
#pragma unroll
            for (int i = 0; i < N; ++i) x[i] = best_x[i];
            R.fnorm = best_fn;
        }
        return R;
    }
};

}  // namespace gs
