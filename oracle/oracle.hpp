// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A C++ restatement of GeometricSolvers.jl's R1 solve, the per-system body of the R2 batched
// solve `solve_batch!` (src/nonlinear/batched.jl:171-179). Each function cites the reference
// file:line it restates. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
// load it, and only as the checker or the timed CPU baseline; the product path never links it.
//
// Build: -O2 -ffp-contract=off. Julia does not contract `a*b + c` into an FMA; the places where the
// reference itself calls `muladd` (StaticArrays' unrolled matrix-vector product) are written as
// explicit `std::fma` here.
//
// Parity status of each part is stated in the header comment of the function; "pinned" means a
// test in tests/ checks it against a golden vector, a known-answer test or a closed form.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <utility>

namespace orc {

enum Code : uint8_t {
    SUCCESS = 0,
    STALLED = 1,
    MAXITERS = 2,
    SINGULAR = 3,
    NONFINITE = 4,
    LINESEARCH_FAILED = 5
};

template <class R> inline R nan_() { return std::numeric_limits<R>::quiet_NaN(); }
template <class R> inline R eps_() { return std::numeric_limits<R>::epsilon(); }
template <class R> inline R floatmin_() { return std::numeric_limits<R>::min(); }
template <class R> inline R floatmax_() { return std::numeric_limits<R>::max(); }
template <class R> inline R tiny_() { return std::numeric_limits<R>::denorm_min(); }
// exponent(floatmax(R)): 1023 for Float64, 127 for Float32.
template <class R> inline int emax_() { return std::numeric_limits<R>::max_exponent - 1; }

template <class T, int N> using Vec = std::array<T, N>;
// Column-major N×N, as an SMatrix: A(i, j) = a[i + j N].
template <class T, int N> using Mat = std::array<T, N * N>;

// ---------------------------------------------------------------------------------------------
// Status, step info and options (src/base/status.jl:101-262, src/base/options.jl:58-307)
// ---------------------------------------------------------------------------------------------

// SolverStatus{R} (status.jl:101-117).
template <class R> struct Status {
    uint8_t code = SUCCESS;
    int16_t f_calls = 0;
    int32_t iterations = 0;
    R fnorm = 0;
    R stepnorm = 0;
    int32_t step_failures = 0;
    bool promoted = false;
    bool stationary = false;
    int16_t jvp_calls = 0;
    int32_t rank = -1;
    R nullnorm = nan_<R>();
    R gradratio = nan_<R>();
    R rate = nan_<R>();
    bool contracted = false;
    bool reused = false;
};

// StepInfo{R} (status.jl:220-234).
template <class R> struct StepInfo {
    uint8_t code;
    R fnorm;
    R stepnorm;
    int32_t failures;
    bool promoted;
    bool limited;
    bool shortened;
};

template <class R>
inline StepInfo<R> step_info(uint8_t code, R fnorm, R stepnorm, int32_t failures, bool promoted,
                             bool limited, bool shortened = false) {
    return StepInfo<R>{code, fnorm, stepnorm, failures, promoted, limited, shortened};
}

// record (status.jl:245-262): one step folded into the status; rate = ‖Δₖ‖/‖Δₖ₋₁‖.
template <class R> inline Status<R> record(const Status<R>& s, const StepInfo<R>& info) {
    Status<R> r = s;
    r.rate = s.stepnorm > R(0) ? info.stepnorm / s.stepnorm : nan_<R>();
    r.code = info.code;
    r.iterations = s.iterations + 1;
    r.fnorm = info.fnorm;
    r.stepnorm = info.stepnorm;
    r.step_failures = s.step_failures + info.failures;
    r.promoted = s.promoted | info.promoted;
    return r;
}

enum NormKind { L2NORM = 0, MAXNORM = 1 };

// Options{R, N} (options.jl:58-71).
template <class R> struct Options {
    R f_abstol = 0;
    R f_reltol = std::sqrt(eps_<R>());
    R x_abstol = 0;
    R x_reltol = 2 * eps_<R>();
    int32_t min_iterations = 0;
    int32_t max_iterations = 100;
    int32_t max_stalls = 2;
    R g_tol = std::sqrt(eps_<R>());
    int norm = L2NORM;
    R kappa = R(1) / R(10);
    bool roundoff = false;
    bool rate_test = true;
};

// Julia's max(a, b) of floats propagates NaN; std::max does not. Used wherever a NaN can reach a max.
template <class R> inline R jmax(R a, R b) {
    if (std::isnan(a) || std::isnan(b)) return nan_<R>();
    return a > b ? a : b;
}
template <class R> inline R jmin(R a, R b) {
    if (std::isnan(a) || std::isnan(b)) return nan_<R>();
    return a < b ? a : b;
}

// residual_converged (options.jl:267-271).
template <class R> inline bool residual_converged(const Options<R>& o, const Status<R>& s, R fnorm0) {
    if (!(s.iterations >= o.min_iterations)) return false;
    return s.fnorm <= jmax(o.f_abstol, o.f_reltol * fnorm0);
}

// step_converged (options.jl:275-280).
template <class R> inline bool step_converged(const Options<R>& o, const Status<R>& s, R xnorm) {
    if (!(s.iterations >= o.min_iterations)) return false;
    return s.iterations > 0 && s.stepnorm > R(0) &&
           s.stepnorm <= jmax(o.x_abstol, o.x_reltol * xnorm);
}

// rate_converged (options.jl:288-296).
template <class R>
inline bool rate_converged(const Options<R>& o, const Status<R>& s, R theta_prev, R xnorm) {
    if (!o.rate_test) return false;
    if (!(s.iterations >= o.min_iterations)) return false;
    R theta = jmax(s.rate, theta_prev);
    if (!(s.stepnorm > R(0) && theta < R(1))) return false;
    return theta / (R(1) - theta) * s.stepnorm <= o.kappa * jmax(o.x_abstol, o.x_reltol * xnorm);
}

// roundoff_stop (options.jl:302-307).
template <class R> inline bool roundoff_stop(const Options<R>& o, const Status<R>& s, R theta_prev) {
    if (!o.roundoff) return false;
    if (!(s.iterations >= o.min_iterations)) return false;
    return s.rate >= R(1) && theta_prev < R(1);
}

// closing_status / end_status / ok_status / contracted_status (src/nonlinear/solver.jl:88-112).
template <class R>
inline Status<R> closing_status(const Status<R>& s, uint8_t code, bool stationary, bool contracted) {
    Status<R> r;
    r.code = code;
    r.f_calls = 0;
    r.iterations = s.iterations;
    r.fnorm = s.fnorm;
    r.stepnorm = s.stepnorm;
    r.step_failures = s.step_failures;
    r.promoted = s.promoted;
    r.stationary = stationary;
    r.jvp_calls = 0;
    r.rank = -1;
    r.nullnorm = nan_<R>();
    r.gradratio = nan_<R>();
    r.rate = s.rate;
    r.contracted = contracted;
    r.reused = false;
    return r;
}

// ---------------------------------------------------------------------------------------------
// Reductions (src/base/reductions.jl:12-85) and power-of-two scaling (src/linear/factors.jl:726-919)
// ---------------------------------------------------------------------------------------------

// ldexp2 (factors.jl:726-755): a·2^e, exact unless the result leaves the range; correctly rounded
// (ties to even) into the subnormal range, ±Inf on overflow, NaN/Inf/0 returned unchanged. This is
// IEEE scalbn in round-to-nearest. Pinned by tests/test_oracle.py (closed forms at the edges).
template <class R> inline R ldexp2(R x, int e) { return std::scalbn(x, e); }

// Julia `exponent` of a finite nonzero value: ⌊log2|x|⌋, also for subnormals.
template <class R> inline int exponent_(R x) { return std::ilogb(x); }

// norm2 (reductions.jl:12-15): Σ|aᵢ|² as a left fold, with no FMA (StaticArrays' unrolled sum).
template <class R, size_t L> inline R norm2(const std::array<R, L>& a) {
    if (L == 0) return R(0);
    R s = a[0] * a[0];
    for (size_t i = 1; i < L; ++i) s = s + a[i] * a[i];
    return s;
}

// rnorm (reductions.jl:41-57): sqrt(norm2) where finite and ≥ length·floatmin, else the scaled path
// with the largest entry's power of two (float sum in at least Float32, as R here).
template <class R, size_t L> inline R rnorm(const std::array<R, L>& a) {
    R s = norm2(a);
    if (std::isfinite(s) && s >= R(L) * floatmin_<R>()) return std::sqrt(s);
    if (L == 0) return R(0);
    R m = std::fabs(a[0]);
    for (size_t i = 1; i < L; ++i) m = jmax(m, std::fabs(a[i]));
    if (m == R(0) || !std::isfinite(m)) return m;
    int e = exponent_(m);
    R t0 = ldexp2(a[0], -e);
    R acc = t0 * t0;
    for (size_t i = 1; i < L; ++i) {
        R t = ldexp2(a[i], -e);
        acc = acc + t * t;
    }
    return ldexp2(std::sqrt(acc), e);
}

// rmaxnorm (reductions.jl:67): max |aᵢ|, NaN-propagating.
template <class R, size_t L> inline R rmaxnorm(const std::array<R, L>& a) {
    if constexpr (L == 0) return R(0);
    R m = std::fabs(a[0]);
    for (size_t i = 1; i < L; ++i) m = jmax(m, std::fabs(a[i]));
    return m;
}

// rdot (reductions.jl:77-85): Σ aᵢ bᵢ as a left fold, no FMA.
template <class R, size_t L> inline R rdot(const std::array<R, L>& a, const std::array<R, L>& b) {
    R s = a[0] * b[0];
    for (size_t i = 1; i < L; ++i) s = s + a[i] * b[i];
    return s;
}

// abs_bits + matrix_exponent (factors.jl:757-759, 826-843): exponent of the largest |entry|,
// clamped to ±emax; 0 for a zero or non-finite array (a NaN has the largest abs bits).
template <class R, size_t L> inline int matrix_exponent(const std::array<R, L>& a) {
    using U = typename std::conditional<sizeof(R) == 8, uint64_t, uint32_t>::type;
    const U signmask = U(1) << (sizeof(R) * 8 - 1);
    U mb = 0;
    for (size_t i = 0; i < L; ++i) {
        U b;
        std::memcpy(&b, &a[i], sizeof(R));
        b &= ~signmask;
        if (b > mb) mb = b;
    }
    R m;
    std::memcpy(&m, &mb, sizeof(R));
    if (!std::isfinite(m) || m == R(0)) return 0;
    int e = exponent_(m);
    int em = emax_<R>();
    return e < -em ? -em : (e > em ? em : e);
}

// headroom (factors.jl:889-890): 2 + ⌈log2 n⌉ bits, for n ≥ 1.
inline int headroom(int n) {
    uint64_t v = uint64_t(n - 1);
    int lz = v == 0 ? 64 : __builtin_clzll(v);
    return 2 + 64 - lz;
}

// rhs_exponent (factors.jl:880-885).
template <class R, size_t L> inline int rhs_exponent(const std::array<R, L>& r, int e) {
    if (L == 0) return 0;
    int em = emax_<R>();
    int m = matrix_exponent(r);
    int lo = std::min(e, 0);
    int c = m < lo ? lo : (m > 0 ? 0 : m);
    return std::max(c, m - (em - headroom(int(L))));
}

// ---------------------------------------------------------------------------------------------
// Static LU (src/linear/static/lu.jl:24-92, src/linear/factors.jl:134-148, 802-811, 916-949)
// ---------------------------------------------------------------------------------------------

template <class TF, int N> struct StaticLU {
    Mat<TF, N> f;
    std::array<int32_t, N> ipiv;
    int32_t info;
    int32_t scale_exponent;
};

// factor_matrix (factors.jl:802-811): J/2^e in TF; in the same type only e < 0 divides.
template <class TF, class T, int N>
inline Mat<TF, N> factor_matrix(const Mat<T, N>& J, int32_t& e_out) {
    int e = matrix_exponent(J);
    constexpr bool same = std::is_same<TF, T>::value;
    Mat<TF, N> A;
    if (e == 0 || (same && !(e < 0))) {
        for (int k = 0; k < N * N; ++k) A[k] = TF(J[k]);
        e_out = 0;
        return A;
    }
    for (int k = 0; k < N * N; ++k) A[k] = TF(ldexp2(J[k], -e));
    e_out = e;
    return A;
}

// factorize!! StaticLU (lu.jl:24-64): first-largest |a| pivot, row swap, a/p multipliers, rank-1
// Schur update without FMA; info = first zero pivot. Pinned by the lu.jl KATs in tests/.
template <class TF, class T, int N> inline StaticLU<TF, N> lu_factorize(const Mat<T, N>& J) {
    StaticLU<TF, N> F;
    F.f = factor_matrix<TF, T, N>(J, F.scale_exponent);
    auto A = [&](int i, int j) -> TF& { return F.f[i + j * N]; };
    F.info = 0;
    for (int k = 0; k < N; ++k) {
        int p = k;
        TF amax = std::fabs(A(k, k));
        for (int i = k + 1; i < N; ++i) {
            TF a = std::fabs(A(i, k));
            if (a > amax) {
                amax = a;
                p = i;
            }
        }
        F.ipiv[k] = int32_t(p + 1);
        if (p != k)
            for (int j = 0; j < N; ++j) std::swap(A(k, j), A(p, j));
        TF akk = A(k, k);
        if (akk == TF(0)) {
            if (F.info == 0) F.info = int32_t(k + 1);
        } else {
            for (int i = k + 1; i < N; ++i) A(i, k) = A(i, k) / akk;
            for (int j = k + 1; j < N; ++j)
                for (int i = k + 1; i < N; ++i) A(i, j) = A(i, j) - A(i, k) * A(k, j);
        }
    }
    return F;
}

// backsolve (lu.jl:67-92): permute, forward unit-L, backward U.
template <class TF, int N> inline Vec<TF, N> lu_backsolve(const StaticLU<TF, N>& F, Vec<TF, N> x) {
    auto A = [&](int i, int j) -> TF { return F.f[i + j * N]; };
    for (int k = 0; k < N; ++k) {
        int p = F.ipiv[k] - 1;
        if (p != k) std::swap(x[k], x[p]);
    }
    for (int k = 0; k < N; ++k) {
        TF xk = x[k];
        for (int i = k + 1; i < N; ++i) x[i] = x[i] - A(i, k) * xk;
    }
    for (int k = N - 1; k >= 0; --k) {
        TF xk = x[k] / A(k, k);
        x[k] = xk;
        for (int i = 0; i < k; ++i) x[i] = x[i] - A(i, k) * xk;
    }
    return x;
}

// ldiv!! with the factors in the element type of r (factors.jl:931-936): scale r by 2^-k,
// backsolve, unscale by 2^(k-e).
template <class T, int N> inline Vec<T, N> ldiv_same(const StaticLU<T, N>& F, const Vec<T, N>& r) {
    int k = rhs_exponent(r, F.scale_exponent);
    Vec<T, N> rh = r;
    if (k != 0)
        for (int i = 0; i < N; ++i) rh[i] = ldexp2(r[i], -k);
    Vec<T, N> d = lu_backsolve<T, N>(F, rh);
    int e = k - F.scale_exponent;
    if (e != 0)
        for (int i = 0; i < N; ++i) d[i] = ldexp2(d[i], e);
    return d;
}

// ldiv!! with factors in TF ≠ T (factors.jl:943-949): r/s with s = max|r|, solve in TF, then
// a·frac·2^(ex-e) in the wider type.
template <class TF, class T, int N>
inline Vec<T, N> ldiv_mixed(const StaticLU<TF, N>& F, const Vec<T, N>& r) {
    T s = std::fabs(r[0]);
    for (int i = 1; i < N; ++i) s = jmax(s, std::fabs(r[i]));
    Vec<T, N> z{};
    if (s == T(0)) return z;
    Vec<TF, N> rt;
    for (int i = 0; i < N; ++i) rt[i] = TF(r[i] / s);
    Vec<TF, N> d = lu_backsolve<TF, N>(F, rt);
    int ex = 0;
    T frac = std::frexp(s, &ex);
    if (!std::isfinite(s)) ex = 0;
    for (int i = 0; i < N; ++i) z[i] = T(ldexp2(T(d[i]) * frac, ex - F.scale_exponent));
    return z;
}

template <class TF, class T, int N> inline Vec<T, N> ldiv(const StaticLU<TF, N>& F, const Vec<T, N>& r) {
    if constexpr (std::is_same<TF, T>::value)
        return ldiv_same<T, N>(F, r);
    else
        return ldiv_mixed<TF, T, N>(F, r);
}

// StaticArrays' SMatrix × SVector (combine_products: a left fold of `muladd`), row i of A x.
template <class T, int N> inline Vec<T, N> matvec(const Mat<T, N>& A, const Vec<T, N>& v) {
    Vec<T, N> y;
    for (int i = 0; i < N; ++i) {
        T acc = A[i] * v[0];
        for (int j = 1; j < N; ++j) acc = std::fma(A[i + j * N], v[j], acc);
        y[i] = acc;
    }
    return y;
}

// adjoint(A) × v in the same form: entry k folds column k of A.
template <class T, int N> inline Vec<T, N> matvec_t(const Mat<T, N>& A, const Vec<T, N>& v) {
    Vec<T, N> y;
    for (int k = 0; k < N; ++k) {
        T acc = A[0 + k * N] * v[0];
        for (int j = 1; j < N; ++j) acc = std::fma(A[j + k * N], v[j], acc);
        y[k] = acc;
    }
    return y;
}

// jvp!! (src/ad/jvp.jl:87-100): a problem with an analytic Jacobian takes J(x) * v; one without
// (the standard test set) defines Jv(x, v), a dual evaluation of F seeded along v.
template <class P, class R, int N, class = void> struct HasJv : std::false_type {};
template <class P, class R, int N>
struct HasJv<P, R, N, std::void_t<decltype(std::declval<P&>().Jv(std::declval<const Vec<R, N>&>(),
                                                                std::declval<const Vec<R, N>&>()))>>
    : std::true_type {};
template <class R, int N, class P> inline Vec<R, N> jvp(P& prob, const Vec<R, N>& x, const Vec<R, N>& v) {
    if constexpr (HasJv<P, R, N>::value) return prob.Jv(x, v);
    else return matvec<R, N>(prob.J(x), v);
}

// Julia's x^n for a Float64 base and an integer n (Base.Math.pow_body, Julia 1.13; a dependency not
// under the reference tree), which x^y with an integer-valued Float64 y runs (deflation.jl's
// d^(-power)): power by squaring of the reciprocal for n < 0, with a compensated running product
// (two_mul = (x·y, fma(x, y, −x·y)), muladd = fma); x = 1 gives 1 exactly.
inline double pow_int(double x, int64_t n) {
    if (x == 1.0 || n == 0) return 1.0;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    if (n < 0) {
        const double rx = 1.0 / x;
        if (n == -1) return rx;
        if (std::isfinite(x)) xnlo = -std::fma(x, rx, -1.0) * rx;
        x = rx;
        n = -n;
    }
    while (n > 1) {
        if (n & 1) {
            const double err = std::fma(y, xnlo, x * ynlo);
            const double hi = x * y;
            ynlo = std::fma(x, y, -hi) + err;
            y = hi;
        }
        const double err = (x * 2.0) * xnlo;
        const double hi = x * x;
        xnlo = std::fma(x, x, -hi) + err;
        x = hi;
        n >>= 1;
    }
    const double err = std::fma(y, xnlo, x * ynlo);
    return (std::isfinite(x) && std::isfinite(err)) ? std::fma(x, y, err) : x * y;
}

// Deflated(prob, roots; power = p, shift = σ) (src/nonlinear/deflation.jl) around problem P: the
// residual M(x) F(x), M = ∏ᵢ (‖x − rᵢ‖^(−p) + σ) (_deflate_scale, :251-261); the Jacobian
// M J + F ∇Mᵀ (:339-343) with (M, ∇M) from the recurrence g ← mᵢ g + M ∇mᵢ, M ← M mᵢ and the
// unit-scaled ∇mᵢ coordinates −(p((uⱼ/d)ϕ)/d), zero where uⱼ = 0 (_deflate_factors, :239-278);
// the product M (J v) + ⟨∇M, v⟩ F (_deflated_jvp, :352-357).
template <class R, int N, class P> struct DeflatedProb {
    static constexpr bool kDeflated = true;
    P in;
    int nroots;
    int64_t power;
    R shift;
    Vec<R, N> roots[4];
    long nf = 0;
    R scale(const Vec<R, N>& x) const {
        R M = R(1);
        for (int r = 0; r < nroots; ++r) {
            Vec<R, N> u;
            for (int i = 0; i < N; ++i) u[i] = x[i] - roots[r][i];
            M = M * (R(pow_int(double(rnorm(u)), -power)) + shift);
        }
        return M;
    }
    void factors(const Vec<R, N>& x, R& M, Vec<R, N>& g) const {
        M = R(1);
        g.fill(R(0));
        const R pw = R(power);
        for (int r = 0; r < nroots; ++r) {
            Vec<R, N> u;
            for (int i = 0; i < N; ++i) u[i] = x[i] - roots[r][i];
            const R d = rnorm(u);
            const R phi = R(pow_int(double(d), -power));
            const R m = phi + shift;
            for (int i = 0; i < N; ++i) {
                const R gi = u[i] == R(0) ? R(0) : -((pw * ((u[i] / d) * phi)) / d);
                g[i] = m * g[i] + M * gi;
            }
            M = M * m;
        }
    }
    Vec<R, N> F(const Vec<R, N>& x) {
        ++nf;
        const R M = scale(x);
        Vec<R, N> f = in.F(x);
        for (int i = 0; i < N; ++i) f[i] = M * f[i];
        return f;
    }
    Mat<R, N> J(const Vec<R, N>& x) {
        Vec<R, N> f = in.F(x);
        Mat<R, N> Jm = in.J(x);
        R M;
        Vec<R, N> g;
        factors(x, M, g);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) Jm[i + j * N] = M * Jm[i + j * N] + f[i] * g[j];
        return Jm;
    }
    Vec<R, N> Jv(const Vec<R, N>& x, const Vec<R, N>& v) {
        Vec<R, N> f = in.F(x);
        Vec<R, N> jv = jvp<R, N>(in, x, v);
        R M;
        Vec<R, N> g;
        factors(x, M, g);
        const R gv = rdot(g, v);
        Vec<R, N> out;
        for (int i = 0; i < N; ++i) out[i] = M * jv[i] + gv * f[i];
        return out;
    }
};

template <class P, class = void> struct is_deflated_prob : std::false_type {};
template <class P> struct is_deflated_prob<P, std::void_t<decltype(P::kDeflated)>> : std::true_type {};

// A test fixture that is a method of its own (the reference's stopping tests define step!! for a
// custom method, test/nonlinear/stopping.jl:553-590): its `custom_step` replaces the Newton step.
template <class P, class = void> struct has_custom_step : std::false_type {};
template <class P>
struct has_custom_step<P, std::void_t<decltype(&P::custom_step)>> : std::true_type {};

// grad_ratio_and_norm (src/nonlinear/newton.jl:1287-1298): ‖Jᴴ F‖/(‖J‖_F ‖F‖) and ‖J‖_F.
template <class T, int N>
inline void grad_ratio_and_norm(const Mat<T, N>& J, const Vec<T, N>& Fx, T fnorm, T& gr, T& jnorm) {
    jnorm = rnorm(J);
    if (jnorm == T(0) || fnorm == T(0)) {
        gr = T(0);
        return;
    }
    T gnorm = rnorm(matvec_t<T, N>(J, Fx));
    if (std::isfinite(gnorm) || !(std::isfinite(jnorm) && std::isfinite(fnorm))) {
        gr = gnorm / jnorm / fnorm;
        return;
    }
    int e = exponent_(fnorm);
    Vec<T, N> Fs;
    T sc = ldexp2(T(1), -e);
    for (int i = 0; i < N; ++i) Fs[i] = Fx[i] * sc;
    T gs = rnorm(matvec_t<T, N>(J, Fs));
    gr = gs / jnorm / ldexp2(fnorm, -e);
}

// regularize!! (src/base/bang.jl:118-127): J + μI, untouched at μ = 0.
template <class T, int N> inline Mat<T, N> regularize(Mat<T, N> J, T mu) {
    if (mu == T(0)) return J;
    for (int i = 0; i < N; ++i) J[i + i * N] = J[i + i * N] + mu;
    return J;
}

// ---------------------------------------------------------------------------------------------
// Method description
// ---------------------------------------------------------------------------------------------

enum Glob { G_STATIC = 0, G_BACKTRACKING = 1, G_LM = 2, G_DOGLEG = 3 };

template <class R> struct Method {
    bool factor_f32 = false;   // LUFactorization(Float32)
    int glob = G_STATIC;
    R static_alpha = 1;
    R bt_c1 = R(1e-4);
    R bt_p = R(0.5);
    int32_t bt_maxiter = 100;
    bool refinement = false;   // Refinement(η, maxiter)
    R ref_eta = std::sqrt(eps_<R>());
    int32_t ref_maxiter = 5;
    R regularization = 0;
    bool best = false;
    int32_t reuse = 0;         // Newton(; reuse = k); 0 = nothing
    R dl_radius = 0;           // DogLeg(; radius), clamped to [1e-6, 1e10]; 0 = nothing
    // LevenbergMarquardt
    R lm_mu0 = R(1e-6);
    R lm_mu_min = std::max(R(1e-15), 4 * eps_<R>());
    R lm_mu_max = R(1e12);
    R lm_mu_decrease = R(1) / R(3);
    R lm_mu_increase = R(4);
    R lm_rho_accept = 0;
    bool lm_diagonal = false;  // LevenbergMarquardt(; scaling = :diagonal)
};

// A problem restated for the oracle: F(x) and J(x) of one system, with its parameters bound, and a
// counter of residual evaluations (the `f_evals` of the reference's golden table).
//   struct P { static constexpr int N; Vec<T,N> F(const Vec<T,N>&); Mat<T,N> J(const Vec<T,N>&); long nf; };

// ---------------------------------------------------------------------------------------------
// Line searches (src/globalization/linesearch.jl:154-258, backtracking.jl:62-132, static.jl:8-30,
// linefunction.jl:114-199)
// ---------------------------------------------------------------------------------------------

template <class R> inline R roundoff(R phi0) {
    // roundoff (linesearch.jl:214)
    return std::max(4 * eps_<R>() * std::fabs(phi0), 4 * tiny_<R>());
}
template <class R> inline R smallest_step(R d0, R tau) {
    // smallest_step (linesearch.jl:223)
    return jmax(tau / std::fabs(d0), floatmin_<R>());
}
template <class R> inline bool sufficient_decrease(R pa, R p0, R demand, R tau) {
    // sufficient_decrease (linesearch.jl:234)
    return pa - p0 <= jmin(R(0), demand + tau);
}
template <class R> inline uint8_t floor_code(R pa, R p0, R tau) {
    // floor_code (linesearch.jl:250)
    return std::fabs(pa - p0) <= tau ? STALLED : LINESEARCH_FAILED;
}
template <class R> inline uint8_t classify(R pa, R p0, R tau) {
    // classify (linesearch.jl:243-246)
    if (std::isfinite(pa) && pa - p0 < -tau) return SUCCESS;
    return floor_code(pa, p0, tau);
}

// backtrack_step (backtracking.jl:320-340).
template <class R> inline R backtrack_step(R phi0, R d0, R a, R pa, R ap, R pp, R p) {
    if (!std::isfinite(pa)) return R(0.1) * a;
    R an = nan_<R>();
    R s = std::fabs(d0);
    R r1 = (pa - phi0) / s + a;
    if (!(ap > R(0))) {
        if (r1 > R(0)) an = (a * a) / (2 * r1);
    } else if (std::isfinite(pp) && a != ap) {
        R q1 = r1 / (a * a);
        R q2 = ((pp - phi0) / s + ap) / (ap * ap);
        R A = (q1 - q2) / (a - ap);
        R B = (a * q2 - ap * q1) / (a - ap);
        R disc = B * B + 3 * A;
        if (disc >= R(0)) an = B > R(0) ? R(1) / (B + std::sqrt(disc)) : (-B + std::sqrt(disc)) / (3 * A);
    }
    if (!(std::isfinite(an) && an > R(0))) an = p * a;
    return jmin(jmax(an, R(0.1) * a), p * a);
}

template <class R, int N> struct LSResult {
    R alpha;
    R phi;
    uint8_t code;
    int32_t evaluations;
    Vec<R, N> residual;
};

// linesearch + search for Backtracking along x + αd (linesearch.jl:154-171, backtracking.jl:354-390).
// `exact` selects ExactStep (slope −2φ₀) or MeasuredSlope (one φ′ at α = 0 via J(x) d).
template <class R, int N, class P>
inline LSResult<R, N> backtracking(P& prob, const Method<R>& m, const Vec<R, N>& x, const Vec<R, N>& Fx,
                                   const Vec<R, N>& d, R phi0, bool exact) {
    LSResult<R, N> res;
    R alpha = R(1);
    R d0;
    int32_t n = 0;
    if (exact) {
        d0 = -2 * phi0;
    } else {
        // φ′(lf, 0) (linefunction.jl:47-52): x + 0·d, F there, J there times d, 2 rdot.
        Vec<R, N> xa;
        for (int i = 0; i < N; ++i) xa[i] = x[i] + R(0) * d[i];
        Vec<R, N> Fa = prob.F(xa);
        Vec<R, N> Jv = jvp<R, N>(prob, xa, d);
        d0 = 2 * rdot(Fa, Jv);
        n = 1;
    }
    if (!(std::isfinite(phi0) && std::isfinite(d0) && d0 < R(0))) {
        res.alpha = alpha;
        res.phi = nan_<R>();
        res.code = !(std::isfinite(phi0) && std::isfinite(d0)) ? NONFINITE
                   : (d0 > R(0) ? LINESEARCH_FAILED : STALLED);
        res.evaluations = n;
        res.residual = Fx;
        return res;
    }
    R tau = roundoff(phi0);
    alpha = std::min(std::max(alpha, smallest_step(d0, tau)), std::numeric_limits<R>::infinity());
    // search (backtracking.jl:354-390)
    R amin = smallest_step(d0, tau);
    R aa = alpha, pa = nan_<R>();
    Vec<R, N> Fa = Fx;
    R ap = nan_<R>(), pp = nan_<R>();
    int frozen = 0;
    uint8_t code = LINESEARCH_FAILED;
    bool done = false;
    int32_t trips = 0;
    const R sqrteps = std::sqrt(eps_<R>());
    while (!done && trips < m.bt_maxiter) {
        trips += 1;
        aa = alpha;
        Vec<R, N> xa;
        for (int i = 0; i < N; ++i) xa[i] = x[i] + alpha * d[i];
        Fa = prob.F(xa);
        pa = norm2(Fa);
        n += 1;
        frozen = (pa == phi0 && alpha <= sqrteps) ? frozen + 1 : 0;
        if (sufficient_decrease(pa, phi0, m.bt_c1 * alpha * d0, tau)) {
            code = classify(pa, phi0, tau);
            done = true;
        } else if (frozen >= 2) {
            code = STALLED;
            done = true;
        } else if (alpha <= amin) {
            code = floor_code(pa, phi0, tau);
            done = true;
        } else {
            R an = std::isfinite(pa) ? backtrack_step(phi0, d0, alpha, pa, ap, pp, m.bt_p)
                                     : jmin(R(0.1) * alpha, std::sqrt(alpha) * std::sqrt(amin));
            ap = alpha;
            pp = pa;
            alpha = jmax(an, amin);
        }
    }
    res.alpha = aa;
    res.phi = pa;
    res.code = code;
    res.evaluations = n;
    res.residual = Fa;
    return res;
}

// ---------------------------------------------------------------------------------------------
// The Newton state of one R1 solve (src/nonlinear/newton.jl:688-744, 1135-1231)
// ---------------------------------------------------------------------------------------------

template <class R, int N> struct NState {
    R merit;        // norm2(Fx) for a line search; unused (nothing) for Static
    R fnorm;        // ‖Fx‖ the state carries
    int32_t rank;   // NewtonDiagnostics (newton.jl:688-697)
    R nullnorm;
    R gradratio;
    R jnorm;
    bool promoted;  // mixed precision: promoted to the working precision
    int32_t slow_steps;
    // reuse = k: the factor the state holds (in the factor precision) and its rule state
    // FactorAge (newton.jl:621-633)
    StaticLU<R, N> facR;
    StaticLU<float, N> facF;
    int32_t age = 0;
    bool stalled = false;
    bool reused = false;
    // DogLeg: the trust radius TrustRadius (newton.jl:637-640); 0 until the first step sets it
    R radius = 0;
    // best iterate (newton.jl:760-799)
    Vec<R, N> best_x;
    R best_fnorm;
};

// The step of one plain Newton iteration in the working precision (newton.jl:1671-1688) or the
// mixed-precision step (newton.jl:1604-1720), with the Static or Backtracking globalization.
template <class R, int N, class P> struct Solver {
    P& prob;
    const Method<R>& m;
    const Options<R>& o;
    Solver(P& p, const Method<R>& mm, const Options<R>& oo) : prob(p), m(mm), o(oo) {}

    // Every residual evaluation goes through the problem, which counts it in `prob.nf`.
    Vec<R, N> F(const Vec<R, N>& x) { return prob.F(x); }

    // residual_norm (newton.jl:1274-1278).
    R residual_norm(const Vec<R, N>& Fx, const NState<R, N>& st) const {
        if (m.glob == G_BACKTRACKING) {
            R merit = st.merit;
            if (std::isfinite(merit) && merit >= R(N) * floatmin_<R>()) return std::sqrt(merit);
            return rnorm(Fx);
        }
        return std::isnan(st.fnorm) ? rnorm(Fx) : st.fnorm;
    }

    // globalize!! Static (linefunction.jl:114-131) and globalize_merit!! for a line search
    // (linefunction.jl:159-199). Returns the step info and updates x, Fx and the merit.
    StepInfo<R> globalize(Vec<R, N>& x, Vec<R, N>& Fx, const Vec<R, N>& d, bool exact, R& merit) {
        if (m.glob == G_STATIC) {
            R s = m.static_alpha;
            R dn = rnorm(d);
            R stepnorm = std::fabs(s) * dn;
            if (!std::isfinite(stepnorm)) return step_info<R>(NONFINITE, rnorm(Fx), stepnorm, 0, false, false);
            Vec<R, N> xp;
            for (int i = 0; i < N; ++i) xp[i] = x[i] + s * d[i];
            Vec<R, N> Fp = F(xp);
            R fnp = rnorm(Fp);
            if (!std::isfinite(fnp)) return step_info<R>(NONFINITE, rnorm(Fx), stepnorm, 0, false, false);
            x = xp;
            Fx = Fp;
            return step_info<R>(SUCCESS, fnp, stepnorm, 0, false, false);
        }
        R phi0 = merit;
        R dn = rnorm(d);
        if (!std::isfinite(dn)) return step_info<R>(NONFINITE, rnorm(Fx), dn, 0, false, false);
        LSResult<R, N> res = backtracking<R, N>(prob, m, x, Fx, d, phi0, exact);
        if (res.code == SUCCESS) {
            for (int i = 0; i < N; ++i) x[i] = x[i] + res.alpha * d[i];
            Fx = res.residual;
            merit = res.phi;
            return step_info<R>(SUCCESS, std::sqrt(res.phi), res.alpha * dn, 0, false, false, res.alpha < R(1));
        } else if (res.code == LINESEARCH_FAILED) {
            if (std::isfinite(res.phi) && res.phi < phi0) {
                for (int i = 0; i < N; ++i) x[i] = x[i] + res.alpha * d[i];
                Fx = res.residual;
                merit = res.phi;
                return step_info<R>(LINESEARCH_FAILED, std::sqrt(res.phi), res.alpha * dn, 1, false, false,
                                    res.alpha < R(1));
            }
            return step_info<R>(LINESEARCH_FAILED, rnorm(Fx), R(0), 1, false, false);
        } else if (res.code == STALLED) {
            return step_info<R>(STALLED, rnorm(Fx), R(0), 0, false, false);
        }
        return step_info<R>(NONFINITE, rnorm(Fx), R(0), 0, false, false);
    }

    // The product a refinement takes (newton.jl:1344-1384): the held regularised Jacobian of a
    // fresh factor (HeldJacobianJVP), or, along a reused factor (`held` = nullptr), the JVP at x,
    // which is RegularizedJVP's 1·(J(x) v) + μ v where μ > 0.
    Vec<R, N> product(const Mat<R, N>* held, const Vec<R, N>& x, const Vec<R, N>& v) {
        if (held) return matvec<R, N>(*held, v);
        Vec<R, N> Jv = jvp<R, N>(prob, x, v);
        if (m.regularization == R(0)) return Jv;
        for (int i = 0; i < N; ++i) Jv[i] = R(1) * Jv[i] + m.regularization * v[i];
        return Jv;
    }

    // refine_passes!! (src/nonlinear/refinement.jl:211-244) with the held Jacobian J.
    template <class TF>
    Vec<R, N> refine(Vec<R, N> d, const StaticLU<TF, N>& fac, const Mat<R, N>& J, const Vec<R, N>& Fx,
                     bool& failed) {
        return refine_with<TF>(d, fac, &J, Fx, Fx, failed);
    }

    // refine_passes!! with the product of `product(held, x, ·)`.
    template <class TF>
    // `fnorm_in` is the caller's ‖F(x)‖ where it holds one (the `fnorm` argument of refine!!); a step
    // passes none (newton_direction, newton.jl:1343-1358), and the norm is formed from Fx.
    Vec<R, N> refine_with(Vec<R, N> d, const StaticLU<TF, N>& fac, const Mat<R, N>* held, const Vec<R, N>& x,
                          const Vec<R, N>& Fx, bool& failed, const R* fnorm_in = nullptr) {
        failed = false;
        if (!m.refinement) return d;
        R fnorm = fnorm_in ? *fnorm_in : rnorm(Fx);
        if (fnorm == R(0)) return d;
        if (!std::isfinite(fnorm)) {
            failed = true;
            return d;
        }
        R eta = std::max(m.ref_eta, std::sqrt(eps_<R>()));
        Vec<R, N> Jv = product(held, x, d);
        Vec<R, N> r;
        for (int i = 0; i < N; ++i) r[i] = -Fx[i] + -Jv[i];
        R rprev = rnorm(r);
        if (!std::isfinite(rprev)) {
            failed = true;
            return d;
        }
        if (rprev <= eta * fnorm) return d;
        for (int32_t it = 1; it <= m.ref_maxiter; ++it) {
            Vec<R, N> dd = ldiv<TF, R, N>(fac, r);
            Vec<R, N> dc;
            for (int i = 0; i < N; ++i) dc[i] = d[i] + dd[i];
            Jv = product(held, x, dc);
            for (int i = 0; i < N; ++i) r[i] = -Fx[i] + -Jv[i];
            R rcur = rnorm(r);
            if (!(rcur < rprev)) {
                failed = true;
                return d;
            }
            d = dc;
            rprev = rcur;
            if (rcur <= eta * fnorm) return d;
        }
        failed = true;
        return d;
    }

    // newton_factor (newton.jl:1302-1311): J, grad ratio, regularize, factorize in TF.
    template <class TF>
    StaticLU<TF, N> factor(const Vec<R, N>& x, const Vec<R, N>& Fx, NState<R, N>& st, Mat<R, N>& Jreg,
                           R& gr, R& jn) {
        R fnorm = residual_norm(Fx, st);
        Mat<R, N> J = prob.J(x);
        grad_ratio_and_norm<R, N>(J, Fx, fnorm, gr, jn);
        Jreg = regularize<R, N>(J, m.regularization);
        return lu_factorize<TF, R, N>(Jreg);
    }

    void set_diag(NState<R, N>& st, R gr, R jn) {
        st.rank = -1;
        st.nullnorm = nan_<R>();
        st.gradratio = gr;
        st.jnorm = jn;
    }

    // newton_step!! plain (newton.jl:1679-1688).
    StepInfo<R> plain_step(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        Mat<R, N> Jreg;
        R gr, jn;
        StaticLU<R, N> fac = factor<R>(x, Fx, st, Jreg, gr, jn);
        StepInfo<R> info;
        R merit = st.merit;
        if (fac.info > 0) {
            info = step_info<R>(SINGULAR, rnorm(Fx), R(0), 0, false, false);
        } else {
            // newton_direction (newton.jl:1344-1349): d = -(J \ F), refined.
            Vec<R, N> d = ldiv<R, R, N>(fac, Fx);
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            bool failed;
            d = refine<R>(d, fac, Jreg, Fx, failed);
            // newton_stepkind (newton.jl:1244-1254): ExactStep iff μ = 0 and no refinement.
            bool exact = m.regularization == R(0) && !m.refinement;
            info = globalize(x, Fx, d, exact, merit);
        }
        set_diag(st, gr, jn);
        st.merit = merit;
        st.fnorm = info.fnorm;
        return info;
    }

    // linear_residual_ratio (newton.jl:1422-1430).
    R linear_residual_ratio(const Mat<R, N>& J, const Vec<R, N>& Fx, const Vec<R, N>& d, R fnorm_old) {
        if (fnorm_old == R(0) || !std::isfinite(fnorm_old)) return R(0);
        Vec<R, N> Jd = matvec<R, N>(J, d);
        Vec<R, N> lin;
        for (int i = 0; i < N; ++i) lin[i] = R(1) * Fx[i] + R(1) * Jd[i];
        return rnorm(lin) / fnorm_old;
    }

    // The result of mixed_newton!! (newton.jl:1471-1478): the direction, the regularised Jacobian
    // the factorisation used, the linear-model ratio, whether the direction exists, the failures.
    struct MixedDir {
        Vec<R, N> d{};
        Mat<R, N> Jreg;
        R lin_rel = 0;
        bool ok = true;
        int32_t failures = 0;
    };

    // mixed_newton!! with a fresh factor and no store (newton.jl:1604-1666): FP32 factor and
    // refinement, promotion to T where either fails; st.promoted and the diagnostics are updated.
    MixedDir mixed_newton(const Vec<R, N>& x, const Vec<R, N>& Fx, NState<R, N>& st, R fnorm_old) {
        MixedDir md;
        Vec<R, N>& d = md.d;
        Mat<R, N>& Jreg = md.Jreg;
        R gr, jn;
        if (st.promoted) {
            StaticLU<R, N> fac = factor<R>(x, Fx, st, Jreg, gr, jn);
            set_diag(st, gr, jn);
            if (fac.info > 0) {
                md.ok = false;
            } else {
                d = ldiv<R, R, N>(fac, Fx);
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                bool failed;
                d = refine<R>(d, fac, Jreg, Fx, failed);
            }
        } else {
            StaticLU<float, N> fac = factor<float>(x, Fx, st, Jreg, gr, jn);
            bool ref_failed = false;
            if (fac.info > 0) {
                ref_failed = true;
            } else {
                d = ldiv<float, R, N>(fac, Fx);
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                d = refine<float>(d, fac, Jreg, Fx, ref_failed);
            }
            set_diag(st, gr, jn);
            if (ref_failed || fac.info > 0) {
                md.failures += 1;
                // promote_factor!! (newton.jl:1332-1339): refactorise Jreg in T.
                StaticLU<R, N> facT = lu_factorize<R, R, N>(Jreg);
                st.promoted = true;
                if (facT.info > 0) {
                    md.ok = false;
                } else {
                    d = ldiv<R, R, N>(facT, Fx);
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                }
            } else {
                md.lin_rel = linear_residual_ratio(Jreg, Fx, d, fnorm_old);
            }
        }
        return md;
    }

    // read_mixed_stall (newton.jl:1437-1451) on the state and the step info.
    void read_mixed_stall(NState<R, N>& st, StepInfo<R>& info, R fnorm_old, R lin_rel) const {
        R eta_lin = std::max(std::sqrt(eps_<R>()), R(1e-4));
        if (info.fnorm > R(0.5) * fnorm_old && lin_rel > eta_lin) {
            int32_t slow = st.slow_steps + 1;
            if (slow >= 2) {
                st.promoted = true;
                st.slow_steps = 0;
                info.failures += 1;
                info.promoted = true;
            } else {
                st.promoted = false;
                st.slow_steps = slow;
            }
        } else {
            st.promoted = false;
            st.slow_steps = 0;
        }
    }

    // mixed_step!! + mixed_newton!! with fresh factors and no reuse (newton.jl:1604-1666, 1692-1720).
    StepInfo<R> mixed_step(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        R fnorm_old = residual_norm(Fx, st);
        bool was_promoted = st.promoted;
        MixedDir md = mixed_newton(x, Fx, st, fnorm_old);
        if (!md.ok) {
            return step_info<R>(SINGULAR, rnorm(Fx), R(0), md.failures, st.promoted, false);
        }
        const bool md_promoted = st.promoted;
        bool promoted_step = md_promoted && !was_promoted;
        // newton_stepkind (newton.jl:1244-1254) with in_working_T = promoted.
        bool exact = m.regularization == R(0) && md_promoted;
        R merit = st.merit;
        StepInfo<R> info = globalize(x, Fx, md.d, exact, merit);
        info.failures += md.failures;  // bump_failures (newton.jl:1480-1485)
        const R lin_rel = md.lin_rel;
        if (promoted_step) {
            info.promoted = true;
            st.promoted = true;
            st.slow_steps = 0;
        } else if (was_promoted) {
            // the promoted state keeps stepping in T
        } else {
            read_mixed_stall(st, info, fnorm_old, lin_rel);
        }
        st.merit = merit;
        st.fnorm = info.fnorm;
        return info;
    }

    // ---- DogLeg (src/globalization/dogleg.jl:116-206, newton.jl:1725-1780) ----

    static R dogleg_lo() { return R(1e-6); }
    static R dogleg_hi() { return R(1e10); }
    static R dogleg_shrink(R d) { return jmax(R(0.25) * d, dogleg_lo()); }
    static R dogleg_grow(R d) { return jmin(R(4) * d, dogleg_hi()); }
    // clamp(δ, lo, hi) in Julia's form: ifelse(δ > hi, hi, ifelse(δ < lo, lo, δ))
    static R dogleg_clamp(R d) { return d > dogleg_hi() ? dogleg_hi() : (d < dogleg_lo() ? dogleg_lo() : d); }

    // dogleg_step (dogleg.jl:116-206): the trial loop on the model matrix J (the regularised
    // Jacobian that was factorised) with the Newton step sN (read where `newton`) and the radius
    // `radius` (0 when unset); x, Fx and the radius are updated.
    StepInfo<R> dogleg_step(Vec<R, N>& x, Vec<R, N>& Fx, const Mat<R, N>& J, const Vec<R, N>& sN, bool newton,
                            R& radius) {
        const R lo = dogleg_lo();
        const R fnorm = rnorm(Fx);
        if (fnorm == R(0)) return step_info<R>(SUCCESS, fnorm, R(0), 0, false, false);
        const R nN = newton ? rnorm(sN) : std::numeric_limits<R>::infinity();
        Vec<R, N> u, gt, gh;
        const R finv = R(1) / fnorm;
        for (int i = 0; i < N; ++i) u[i] = finv * Fx[i];
        gt = matvec_t<R, N>(J, u);
        const R gnorm = rnorm(gt);
        const R ginv = R(1) / gnorm;
        for (int i = 0; i < N; ++i) gh[i] = ginv * gt[i];
        const Vec<R, N> Jg = matvec<R, N>(J, gh);
        const R jnorm = rnorm(Jg);
        const bool cauchy = gnorm != R(0) && std::isfinite(gnorm) && jnorm != R(0) && std::isfinite(jnorm);
        const R nC = fnorm * ((gnorm / jnorm) / jnorm);
        R delta = radius;
        if (delta == R(0) && (newton || cauchy)) {
            const R n0 = newton ? nN : nC;
            delta = n0 < lo ? R(2) * lo : dogleg_clamp(n0);
        }
        while (true) {
            Vec<R, N> sv;
            bool boundary;
            if (newton && nN <= delta) {
                sv = sN;
                boundary = nN >= delta;
            } else if (!cauchy) {
                radius = delta;
                return step_info<R>(STALLED, fnorm, R(0), 0, false, false);
            } else if (!newton || nC >= delta) {
                const R a = -jmin(nC, delta);
                for (int i = 0; i < N; ++i) sv[i] = a * gh[i];
                boundary = nC >= delta;
            } else {
                Vec<R, N> av, b, bh;
                const R ac = -(nC / delta);
                for (int i = 0; i < N; ++i) av[i] = ac * gh[i];
                for (int i = 0; i < N; ++i) b[i] = R(1) * sN[i] + nC * gh[i];
                const R binv = R(1) / rnorm(b);
                for (int i = 0; i < N; ++i) bh[i] = binv * b[i];
                const R ab = rdot(av, bh);
                const R c = R(1) - norm2(av);
                const R ab2 = ab * ab;
                const R t = c / (ab + std::sqrt(ab2 + c));
                const R dt = delta * t;
                for (int i = 0; i < N; ++i) sv[i] = delta * av[i] + dt * bh[i];
                boundary = true;
            }
            const Vec<R, N> Js = matvec<R, N>(J, sv);
            Vec<R, N> Flin;
            for (int i = 0; i < N; ++i) Flin[i] = R(1) * Fx[i] + R(1) * Js[i];
            const R pred = fnorm - rnorm(Flin);
            Vec<R, N> xp;
            for (int i = 0; i < N; ++i) xp[i] = x[i] + R(1) * sv[i];
            const Vec<R, N> Fp = F(xp);
            const R fnp = rnorm(Fp);
            const R ared = fnorm - fnp;
            if (pred > R(0) && ared >= R(1e-4) * pred) {
                const bool limited = !(newton && nN <= delta);
                const R rho = ared / pred;
                if (rho < R(0.1)) {
                    delta = nN < delta ? jmax(nN, lo) : dogleg_shrink(delta);
                } else if (rho > R(0.75) && boundary) {
                    delta = dogleg_grow(delta);
                }
                const R stepnorm = rnorm(sv);
                x = xp;
                Fx = Fp;
                radius = delta;
                return step_info<R>(SUCCESS, fnp, stepnorm, 0, false, limited);
            }
            if (delta == lo) {
                radius = delta;
                return step_info<R>(LINESEARCH_FAILED, fnorm, R(0), 1, false, false);
            }
            delta = dogleg_shrink(delta);
        }
    }

    // newton_step!! with DogLeg in the working precision (newton.jl:1725-1737) and
    // dogleg_direction (newton.jl:1741-1746): a failed factor, or a Newton step of zero or
    // non-finite norm, leaves the Cauchy step.
    StepInfo<R> dogleg_plain_step(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        Mat<R, N> Jreg;
        R gr, jn;
        StaticLU<R, N> fac = factor<R>(x, Fx, st, Jreg, gr, jn);
        Vec<R, N> d = Fx;
        bool newton = false;
        if (fac.info <= 0) {
            d = ldiv<R, R, N>(fac, Fx);
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            bool failed;
            d = refine<R>(d, fac, Jreg, Fx, failed);
            const R nd = rnorm(d);
            newton = std::isfinite(nd) && nd != R(0);
        }
        StepInfo<R> info = dogleg_step(x, Fx, Jreg, d, newton, st.radius);
        set_diag(st, gr, jn);
        st.fnorm = info.fnorm;
        return info;
    }

    // mixed_step!! with DogLeg (newton.jl:1757-1780): the direction of mixed_newton!!, the trial loop
    // on the Jacobian it factorised, and the mixed-precision reading of the step.
    StepInfo<R> dogleg_mixed_step(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        const R fnorm_old = residual_norm(Fx, st);
        const bool was_promoted = st.promoted;
        MixedDir md = mixed_newton(x, Fx, st, fnorm_old);
        const bool promoted_step = st.promoted && !was_promoted;
        Vec<R, N> d = Fx;
        bool newton = false;
        if (md.ok) {
            d = md.d;
            const R nd = rnorm(md.d);
            newton = std::isfinite(nd) && nd != R(0);
        }
        StepInfo<R> info = dogleg_step(x, Fx, md.Jreg, d, newton, st.radius);
        info.failures += md.failures;
        if (promoted_step) {
            info.promoted = true;
            st.promoted = true;
            st.slow_steps = 0;
        } else if (!was_promoted) {
            read_mixed_stall(st, info, fnorm_old, md.lin_rel);
        }
        st.fnorm = info.fnorm;
        return info;
    }

    // ---- reuse = k (newton.jl:1781-1874) ----

    // fresh_age / next_age (newton.jl:629-633).
    int32_t fresh_age() const { return m.reuse == 1 ? 0 : 1; }
    int32_t next_age(const NState<R, N>& st) const { return st.age + 1 >= m.reuse ? 0 : st.age + 1; }

    // reuse_result (newton.jl:1814-1820): the rule state after a step along a reused or fresh factor.
    void set_rule(NState<R, N>& st, uint8_t code, int32_t age_next, bool reused) const {
        st.age = code == SINGULAR ? 0 : age_next;
        st.stalled = code == SINGULAR ? false : code == STALLED;
        st.reused = reused;
    }

    // newton_step!! with reuse in the working precision (newton.jl:1781-1810). A step along the held
    // factor refines with the JVP at x; a failed refinement (one step-rule failure) or a line search
    // that fails without a step falls back to a fresh factor within the step.
    StepInfo<R> plain_step_reuse(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        int32_t failures = 0;
        if (!(st.age == 0 || st.stalled)) {
            const int32_t age_next = next_age(st);
            Vec<R, N> d = ldiv<R, R, N>(st.facR, Fx);
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            bool ref_failed = false;
            if (m.refinement) d = refine_with<R>(d, st.facR, nullptr, x, Fx, ref_failed);
            if (ref_failed) {
                failures = 1;
            } else {
                Vec<R, N> x1 = x, F1 = Fx;
                R merit = st.merit;
                // newton_line with fresh = false: MeasuredSlope (newton.jl:1406-1415, 1244-1254)
                StepInfo<R> info = globalize(x1, F1, d, false, merit);
                if (!(info.code == LINESEARCH_FAILED && info.stepnorm == R(0))) {
                    x = x1;
                    Fx = F1;
                    st.merit = merit;
                    st.fnorm = info.fnorm;
                    set_rule(st, info.code, age_next, true);
                    return info;
                }
            }
        }
        Mat<R, N> Jreg;
        R gr, jn;
        StaticLU<R, N> fac = factor<R>(x, Fx, st, Jreg, gr, jn);
        StepInfo<R> info;
        R merit = st.merit;
        if (fac.info > 0) {
            info = step_info<R>(SINGULAR, rnorm(Fx), R(0), 0, false, false);
        } else {
            Vec<R, N> d = ldiv<R, R, N>(fac, Fx);
            for (int i = 0; i < N; ++i) d[i] = -d[i];
            bool failed;
            d = refine<R>(d, fac, Jreg, Fx, failed);
            bool exact = m.regularization == R(0) && !m.refinement;
            info = globalize(x, Fx, d, exact, merit);
        }
        info.failures += failures;  // bump_failures (newton.jl:1480-1485)
        set_diag(st, gr, jn);
        st.facR = fac;
        st.merit = merit;
        st.fnorm = info.fnorm;
        set_rule(st, info.code, fresh_age(), false);
        return info;
    }

    // mixed_step!! with reuse (newton.jl:1823-1874) and mixed_newton!! with store = true
    // (newton.jl:1604-1666): a promoted state factorises in T at every step; otherwise the step runs
    // along the held FP32 factor until its age or a stall refreshes it, and a reused direction whose
    // refinement fails is refactorised in FP32 at x (one failure) before any promotion.
    StepInfo<R> mixed_step_reuse(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        const R fnorm_old = residual_norm(Fx, st);
        const bool state_promoted = st.promoted;
        bool fresh = st.promoted || st.age == 0 || st.stalled;
        const int32_t age_fresh = fresh_age();
        int32_t age_next = fresh ? age_fresh : next_age(st);
        while (true) {
            const bool was_fresh = fresh;
            // --- mixed_newton!! ---
            Vec<R, N> d{};
            bool ok = true, md_fresh = true;
            int32_t failures = 0;
            R lin_rel = 0;
            bool have = false;
            Mat<R, N> Jreg;
            R gr, jn;
            if (st.promoted) {
                StaticLU<R, N> fac = factor<R>(x, Fx, st, Jreg, gr, jn);
                set_diag(st, gr, jn);
                if (fac.info > 0) {
                    ok = false;
                } else {
                    d = ldiv<R, R, N>(fac, Fx);
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                    bool failed;
                    d = refine<R>(d, fac, Jreg, Fx, failed);
                }
                have = true;
            } else if (!fresh) {
                d = ldiv<float, R, N>(st.facF, Fx);
                for (int i = 0; i < N; ++i) d[i] = -d[i];
                bool ref_failed = false;
                if (m.refinement)
                    d = refine_with<float>(d, st.facF, nullptr, x, Fx, ref_failed);
                if (!ref_failed) {
                    md_fresh = false;
                    have = true;
                }
            }
            if (!have) {
                StaticLU<float, N> fac = factor<float>(x, Fx, st, Jreg, gr, jn);
                if (!fresh) failures += 1;
                bool ref_failed = false;
                if (fac.info > 0) {
                    ref_failed = true;
                } else {
                    d = ldiv<float, R, N>(fac, Fx);
                    for (int i = 0; i < N; ++i) d[i] = -d[i];
                    d = refine<float>(d, fac, Jreg, Fx, ref_failed);
                }
                set_diag(st, gr, jn);
                if (ref_failed || fac.info > 0) {
                    failures += 1;
                    // promote_factor!! (newton.jl:1332-1339): refactorise Jreg in T
                    StaticLU<R, N> facT = lu_factorize<R, R, N>(Jreg);
                    st.promoted = true;
                    if (facT.info > 0) {
                        ok = false;
                    } else {
                        d = ldiv<R, R, N>(facT, Fx);
                        for (int i = 0; i < N; ++i) d[i] = -d[i];
                    }
                } else {
                    st.facF = fac;
                    lin_rel = linear_residual_ratio(Jreg, Fx, d, fnorm_old);
                }
            }
            fresh = md_fresh;
            if (!ok) {
                StepInfo<R> info = step_info<R>(SINGULAR, rnorm(Fx), R(0), failures, st.promoted, false);
                set_rule(st, SINGULAR, 0, !md_fresh);
                return info;
            }
            const bool promoted_step = st.promoted && !state_promoted;
            if (promoted_step) {
                age_next = 0;
            } else if (md_fresh && !was_fresh) {
                age_next = age_fresh;
            }
            // newton_stepkind(prob, m, T, md.fresh, promoted) (newton.jl:1244-1254)
            const bool exact = md_fresh && m.regularization == R(0) && st.promoted;
            Vec<R, N> x1 = x, F1 = Fx;
            R merit = st.merit;
            StepInfo<R> info = globalize(x1, F1, d, exact, merit);
            if (!md_fresh && info.code == LINESEARCH_FAILED && info.stepnorm == R(0)) {
                fresh = true;
                age_next = age_fresh;
                continue;
            }
            x = x1;
            Fx = F1;
            info.failures += failures;
            if (promoted_step) {
                info.promoted = true;
                st.promoted = true;
                st.slow_steps = 0;
            } else if (md_fresh && !st.promoted) {
                // read_mixed_stall (newton.jl:1437-1451)
                R eta_lin = std::max(std::sqrt(eps_<R>()), R(1e-4));
                if (info.fnorm > R(0.5) * fnorm_old && lin_rel > eta_lin) {
                    int32_t slow = st.slow_steps + 1;
                    if (slow >= 2) {
                        st.promoted = true;
                        st.slow_steps = 0;
                        info.failures += 1;
                        info.promoted = true;
                    } else {
                        st.slow_steps = slow;
                    }
                } else {
                    st.slow_steps = 0;
                }
            }
            st.merit = merit;
            st.fnorm = info.fnorm;
            set_rule(st, info.code, age_next, !md_fresh);
            return info;
        }
    }

    StepInfo<R> step(Vec<R, N>& x, Vec<R, N>& Fx, NState<R, N>& st) {
        if constexpr (has_custom_step<P>::value) return prob.custom_step(x, Fx);
        StepInfo<R> info = m.glob == G_DOGLEG
                               ? (m.factor_f32 ? dogleg_mixed_step(x, Fx, st) : dogleg_plain_step(x, Fx, st))
                           : m.reuse > 0 ? (m.factor_f32 ? mixed_step_reuse(x, Fx, st) : plain_step_reuse(x, Fx, st))
                                         : (m.factor_f32 ? mixed_step(x, Fx, st) : plain_step(x, Fx, st));
        // step!! keeps the best iterate (newton.jl:1671-1675, take_best :770-773)
        if (m.best && info.fnorm < st.best_fnorm) {
            st.best_x = x;
            st.best_fnorm = info.fnorm;
        }
        return info;
    }

    // finish!! (newton.jl:1165-1231): diagnostics into the status; best restore.
    Status<R> finish(Vec<R, N>& x, const NState<R, N>& st, Status<R> s) {
        Status<R> r = s;
        r.f_calls = 0;
        r.jvp_calls = 0;
        r.rank = st.rank;
        r.nullnorm = st.nullnorm;
        r.gradratio = st.gradratio;
        // reused_factor (newton.jl:1188-1191)
        r.reused = m.reuse > 0 && st.reused && s.iterations > 0;
        // restore_best / should_restore_best (newton.jl:792-799, composition/anderson.jl:503-505)
        if (m.best && r.code != SUCCESS && std::isfinite(st.best_fnorm) && st.best_fnorm < r.fnorm) {
            x = st.best_x;
            r.fnorm = st.best_fnorm;
        }
        return r;
    }

    // solve!! (src/nonlinear/solver.jl:277-378) for a square NonlinearProblem.
    // newton_start!! (newton.jl:1135-1144) and reset_rule (newton.jl:675-686).
    NState<R, N> start(const Vec<R, N>& x, const Vec<R, N>& Fx, R fnorm0) const {
        NState<R, N> st;
        st.merit = m.glob == G_BACKTRACKING ? norm2(Fx) : nan_<R>();
        st.fnorm = fnorm0;
        st.promoted = false;
        st.slow_steps = 0;
        st.rank = -1;
        st.nullnorm = nan_<R>();
        st.gradratio = nan_<R>();
        st.jnorm = 0;
        st.best_x = x;
        st.best_fnorm = fnorm0;
        st.radius = m.dl_radius;  // the radius each solve starts from
        return st;
    }

    // converged_status (solver.jl:122-125): a Deflated problem reports its inner ‖F‖ = ‖M F‖ / M where
    // the norm is finite (_inner_status, deflation.jl:514-528).
    R status_fnorm(const Vec<R, N>& x, R fn) const {
        if constexpr (is_deflated_prob<P>::value) return std::isfinite(fn) ? fn / prob.scale(x) : fn;
        else return fn;
    }
    // test_status (solver.jl:138-146; deflation.jl:540-548 under MaxNorm).
    R test_fnorm(const Vec<R, N>& Fx, R raw, R st) const {
        if (o.norm != MAXNORM) return st;
        if constexpr (is_deflated_prob<P>::value) {
            const R mx = rmaxnorm(Fx);
            return raw == R(0) ? mx : (mx / raw) * st;
        } else {
            return rmaxnorm(Fx);
        }
    }

    Status<R> solve(Vec<R, N>& x) {
        Vec<R, N> Fx = F(x);
        R fnorm0 = rnorm(Fx);
        NState<R, N> st = start(x, Fx, fnorm0);
        Status<R> status0;
        status0.code = SUCCESS;
        status0.iterations = 0;
        status0.fnorm = fnorm0;
        status0.stepnorm = 0;
        if (!std::isfinite(fnorm0)) return finish(x, st, closing_status(status0, NONFINITE, false, false));
        Status<R> status = status0;
        status.fnorm = status_fnorm(x, fnorm0);
        R ftest0 = test_fnorm(Fx, fnorm0, status.fnorm);
        {
            Status<R> tested = status;
            tested.fnorm = ftest0;
            if (residual_converged(o, tested, ftest0)) return finish(x, st, status);
        }
        if (o.max_iterations == 0) return finish(x, st, closing_status(status, MAXITERS, false, false));
        int32_t stalls = 0;
        while (true) {
            Status<R> status_k = status;
            StepInfo<R> info = step(x, Fx, st);
            Status<R> recorded = record(status, info);
            status = recorded;
            status.fnorm = status_fnorm(x, recorded.fnorm);
            // observe_rate!! (newton.jl:1205-1210): a step of rate θ ≥ 1 along a reused factor
            // refreshes the factor at the next step
            if (m.reuse > 0 && st.reused && recorded.rate >= R(1)) st.stalled = true;
            if (info.code == SINGULAR || info.code == NONFINITE || info.code == MAXITERS)
                return finish(x, st, status);
            if (info.code == LINESEARCH_FAILED && info.stepnorm == R(0)) return finish(x, st, status);
            Status<R> tested = status;
            tested.fnorm = test_fnorm(Fx, recorded.fnorm, status.fnorm);
            if (residual_converged(o, tested, ftest0))
                return finish(x, st, closing_status(status, SUCCESS, false, false));
            bool steptested = !(info.code == STALLED || info.limited);
            R xnorm = steptested ? rnorm(x) : nan_<R>();
            bool contracted = !info.shortened && rate_converged(o, status, status_k.rate, xnorm);
            if (contracted ||
                (steptested && (step_converged(o, status, xnorm) ||
                                (info.code == SUCCESS && info.stepnorm == R(0) &&
                                 status.iterations >= o.min_iterations) ||
                                roundoff_stop(o, status, status_k.rate)))) {
                if (contracted) return finish(x, st, closing_status(status, SUCCESS, false, true));
                return finish(x, st, closing_status(status, STALLED, false, false));
            }
            if (info.code == STALLED) {
                stalls += 1;
                if (stalls >= o.max_stalls) return finish(x, st, closing_status(status, STALLED, false, false));
            } else {
                stalls = 0;
            }
            if (status.iterations >= o.max_iterations)
                return finish(x, st, closing_status(status, MAXITERS, false, false));
        }
    }
};

}  // namespace orc
