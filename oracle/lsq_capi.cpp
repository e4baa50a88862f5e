// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// C entry points of the least-squares oracle (lsq.hpp): the exponential-sum fits of the
// least-squares configuration (per-fit R1 `solve(LeastSquaresProblem(F, J), x₀,
// Newton(; globalization = LevenbergMarquardt()), p)`), and the Moré–Garbow–Hillstrom problems of
// the reference's least-squares tests (test/nonlinear/leastsquares.jl:24-103, 601-740) for the KATs.
#include <omp.h>

#include "../include/gsolve.h"
#include "lsq.hpp"

namespace orc {

// The exponential of the exponential-sum family. The family is defined with this fixed algorithm
// (not the platform's exp) so that its F and J have the same bits on every platform that rounds
// IEEE operations correctly:
//   k = rint(x · (1/ln 2)),  r = fma(−k, L2, fma(−k, L1, x))  with L1 = ln 2 rounded to a multiple
//   of 2^-41 (k·L1 exact for |k| < 2^12) and L2 = ln 2 − L1 rounded to double;
//   e^r = Σ_{j≤13} r^j / j! in Horner form with fma, coefficients 1/j! rounded to double;
//   e^x = ldexp(e^r, k); +Inf above ln(2^1024), 0 below ln(2^-1075), NaN unchanged.
// Pinned by tests/test_oracle.py against the C library's exp (≤ 2 ulp apart, almost always equal).
double exp_family(double x) {
    if (std::isnan(x)) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double L1 = 0.6931471805601177, L2 = -1.7239444525614835e-13, INV_LN2 = 1.4426950408889634;
    const double kd = std::rint(x * INV_LN2);
    double r = std::fma(-kd, L1, x);
    r = std::fma(-kd, L2, r);
    // c[j] = 1/j! rounded to double (decimal literals that round-trip to those doubles)
    const double c[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333,
                          0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05,
                          2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08,
                          2.08767569878681e-09, 1.6059043836821613e-10};
    double p = c[13];
    for (int j = 12; j >= 0; --j) p = std::fma(p, r, c[j]);
    return std::ldexp(p, int(kd));
}

// The exponential-sum fit y(t) ≈ Σ_k a_k exp(−λ_k t) with x = (a_1..a_K, λ_1..λ_K), t_i = 5(i−1)/(m−1):
//   F_i = (Σ_k a_k exp(−λ_k t_i)) − y_i  (a left fold over k),
//   J[i, k] = exp(−λ_k t_i),  J[i, K+k] = ((−a_k) t_i) exp(−λ_k t_i),  exp = exp_family.
template <int N> struct ExpSum {
    static constexpr int K = N / 2;
    int mm;
    const double* y;
    long nf = 0;
    int m() const { return mm; }
    double t(int i) const { return (5.0 * double(i)) / double(mm - 1); }
    void F(const double* x, double* r) {
        ++nf;
        for (int i = 0; i < mm; ++i) {
            double ti = t(i);
            double s = x[0] * exp_family(-x[K] * ti);
            for (int k = 1; k < K; ++k) s = s + x[k] * exp_family(-x[K + k] * ti);
            r[i] = s - y[i];
        }
    }
    void J(const double* x, double* Jc) const {
        for (int i = 0; i < mm; ++i) {
            double ti = t(i);
            for (int k = 0; k < K; ++k) {
                double e = exp_family(-x[K + k] * ti);
                Jc[i + size_t(k) * mm] = e;
                Jc[i + size_t(K + k) * mm] = ((-x[k]) * ti) * e;
            }
        }
    }
};

// MGH #11 full-rank linear (m = 10, n = 5), #12 rank-1 linear (10, 5), #8 Bard (15, 3),
// #6 Jennrich–Sampson (10, 2), with their analytic Jacobians.
template <int N> struct Mgh {
    int id;
    long nf = 0;
    int m() const { return id == 8 ? 15 : 10; }
    void F(const double* x, double* r) {
        ++nf;
        if (id == 11) {
            double sx = x[0];
            for (int k = 1; k < 5; ++k) sx = sx + x[k];
            double s = 2.0 / 10.0 * sx;
            for (int i = 0; i < 5; ++i) r[i] = (x[i] - s) - 1.0;
            for (int i = 5; i < 10; ++i) r[i] = -s - 1.0;
        } else if (id == 12) {
            double u = (((x[0] + 2.0 * x[1]) + 3.0 * x[2]) + 4.0 * x[3]) + 5.0 * x[4];
            for (int i = 0; i < 10; ++i) r[i] = double(i + 1) * u - 1.0;
        } else if (id == 8) {
            static const double Y[15] = {0.14, 0.18, 0.22, 0.25, 0.29, 0.32, 0.35, 0.39,
                                         0.37, 0.58, 0.73, 0.96, 1.34, 2.10, 4.39};
            for (int i = 1; i <= 15; ++i) {
                double u = i, v = 16 - i, w = std::min(i, 16 - i);
                r[i - 1] = Y[i - 1] - (x[0] + u / (v * x[1] + w * x[2]));
            }
        } else {
            for (int i = 1; i <= 10; ++i) {
                double ti = i;
                r[i - 1] = (2.0 + 2.0 * ti) - (std::exp(ti * x[0]) + std::exp(ti * x[1]));
            }
        }
    }
    void J(const double* x, double* Jc) const {
        int mm = m();
        if (id == 11) {
            for (int i = 0; i < mm; ++i)
                for (int k = 0; k < 5; ++k) Jc[i + k * mm] = (i == k ? 1.0 : 0.0) - 2.0 / 10.0;
        } else if (id == 12) {
            for (int i = 0; i < mm; ++i)
                for (int k = 0; k < 5; ++k) Jc[i + k * mm] = double(i + 1) * double(k + 1);
        } else if (id == 8) {
            for (int i = 1; i <= 15; ++i) {
                double u = i, v = 16 - i, w = std::min(i, 16 - i);
                double den = v * x[1] + w * x[2];
                Jc[(i - 1) + 0 * mm] = -1.0;
                Jc[(i - 1) + 1 * mm] = u * v / (den * den);
                Jc[(i - 1) + 2 * mm] = u * w / (den * den);
            }
        } else {
            for (int i = 1; i <= 10; ++i) {
                double ti = i;
                Jc[(i - 1) + 0 * mm] = -ti * std::exp(ti * x[0]);
                Jc[(i - 1) + 1 * mm] = -ti * std::exp(ti * x[1]);
            }
        }
    }
};

template <class R> Options<R> options_from_lsq(const gs_options_f64& g) {
    Options<R> o;
    o.f_abstol = g.f_abstol;
    o.f_reltol = g.f_reltol;
    o.x_abstol = g.x_abstol;
    o.x_reltol = g.x_reltol;
    o.min_iterations = g.min_iterations;
    o.max_iterations = g.max_iterations;
    o.max_stalls = g.max_stalls;
    o.g_tol = g.g_tol;
    o.norm = g.norm;
    o.kappa = g.kappa;
    o.roundoff = g.roundoff != 0;
    o.rate_test = g.rate_test != 0;
    return o;
}

Method<double> lm_method_from(const gs_method& g) {
    Method<double> m;
    m.glob = G_LM;
    m.lm_mu0 = g.lm_mu0;
    m.lm_mu_min = g.lm_mu_min;
    m.lm_mu_max = g.lm_mu_max;
    m.lm_mu_decrease = g.lm_mu_decrease;
    m.lm_mu_increase = g.lm_mu_increase;
    m.lm_rho_accept = g.lm_rho_accept;
    m.lm_diagonal = g.lm_scaling == GS_LM_SCALING_DIAGONAL;
    return m;
}

void write_status_lsq(gs_status_f64* out, const Status<double>& s) {
    gs_status_f64 t;
    std::memset(&t, 0, sizeof(t));
    t.code = s.code;
    t.iterations = s.iterations;
    t.fnorm = s.fnorm;
    t.stepnorm = s.stepnorm;
    t.step_failures = s.step_failures;
    t.promoted = s.promoted;
    t.stationary = s.stationary;
    t.rank = s.rank;
    t.nullnorm = s.nullnorm;
    t.gradratio = s.gradratio;
    t.rate = s.rate;
    t.contracted = s.contracted;
    *out = t;
}

template <int N>
void run_expsum(const Method<double>& m, const Options<double>& o, double* x8, int mm, const double* y,
                gs_status_f64* S) {
    ExpSum<N> prob{mm, y};
    LsqSolver<N, ExpSum<N>> s(prob, m, o);
    Vec<double, N> x;
    for (int k = 0; k < N; ++k) x[k] = x8[k];
    Status<double> st = s.solve(x);
    for (int k = 0; k < N; ++k) x8[k] = x[k];
    write_status_lsq(S, st);
}

// F(x) = A x − b (A m×N column-major), with its constant Jacobian: the linear problems of the
// reference's LevenbergMarquardt closed-form tests (test/globalization/levenbergmarquardt.jl:
// 280-330).
template <int N> struct Linear {
    int mm;
    const double* A;
    const double* b;
    long nf = 0;
    int m() const { return mm; }
    void F(const double* x, double* r) {
        ++nf;
        for (int i = 0; i < mm; ++i) {
            double s = A[i] * x[0];
            for (int k = 1; k < N; ++k) s = std::fma(A[i + size_t(k) * mm], x[k], s);
            r[i] = s - b[i];
        }
    }
    void J(const double* x, double* Jc) const {
        (void)x;
        for (size_t i = 0; i < size_t(mm) * N; ++i) Jc[i] = A[i];
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

// The exponential-sum batch: fit i has K[i] ∈ 1..4 terms (n = 2K unknowns, X row i holds x in its
// first n of 8 slots) and m[i] ≥ n points, its data y at Y + y_offset[i].
int orc_solve_expsum_batch(const gs_method* method, const gs_options_f64* options, int64_t nfit, double* X,
                           const int32_t* K, const int32_t* m, const int64_t* y_offset, const double* Y,
                           gs_status_f64* S, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<double> mt = lm_method_from(*method);
    Options<double> o = options_from_lsq<double>(*options);
    int bad = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : bad)
    for (int64_t i = 0; i < nfit; ++i) {
        double* x = X + 8 * i;
        const double* y = Y + y_offset[i];
        switch (K[i]) {
            case 1: run_expsum<2>(mt, o, x, m[i], y, S + i); break;
            case 2: run_expsum<4>(mt, o, x, m[i], y, S + i); break;
            case 3: run_expsum<6>(mt, o, x, m[i], y, S + i); break;
            case 4: run_expsum<8>(mt, o, x, m[i], y, S + i); break;
            default: bad += 1;
        }
    }
    return bad ? GS_EINVAL : GS_OK;
}

// One MGH least-squares problem (11, 12, 8, 6) from x (length n), Newton + LevenbergMarquardt.
int orc_solve_mgh(int id, const gs_method* method, const gs_options_f64* options, double* x, gs_status_f64* S,
                  long* f_evals) {
    Method<double> mt = lm_method_from(*method);
    Options<double> o = options_from_lsq<double>(*options);
    auto run = [&](auto tag) {
        constexpr int N = decltype(tag)::value;
        Mgh<N> prob{id};
        LsqSolver<N, Mgh<N>> s(prob, mt, o);
        Vec<double, N> v;
        for (int k = 0; k < N; ++k) v[k] = x[k];
        Status<double> st = s.solve(v);
        for (int k = 0; k < N; ++k) x[k] = v[k];
        write_status_lsq(S, st);
        if (f_evals) *f_evals = prob.nf;
        return 0;
    };
    switch (id) {
        case 11:
        case 12: return run(std::integral_constant<int, 5>{});
        case 8: return run(std::integral_constant<int, 3>{});
        case 6: return run(std::integral_constant<int, 2>{});
    }
    return GS_EINVAL;
}

// The linear problem A x = b (n ≤ 4 unknowns) solved from x by Newton + LevenbergMarquardt.
int orc_solve_linear_lsq(int n, int mm, const double* A, const double* b, const gs_method* method,
                         const gs_options_f64* options, double* x, gs_status_f64* S) {
    Method<double> mt = lm_method_from(*method);
    Options<double> o = options_from_lsq<double>(*options);
    auto run = [&](auto tag) {
        constexpr int N = decltype(tag)::value;
        Linear<N> prob{mm, A, b};
        LsqSolver<N, Linear<N>> s(prob, mt, o);
        Vec<double, N> v;
        for (int k = 0; k < N; ++k) v[k] = x[k];
        Status<double> st = s.solve(v);
        for (int k = 0; k < N; ++k) x[k] = v[k];
        write_status_lsq(S, st);
        return 0;
    };
    if (mm < n) return GS_EINVAL;
    switch (n) {
        case 1: return run(std::integral_constant<int, 1>{});
        case 2: return run(std::integral_constant<int, 2>{});
        case 3: return run(std::integral_constant<int, 3>{});
        case 4: return run(std::integral_constant<int, 4>{});
    }
    return GS_EINVAL;
}

// lm_damping_diagonal (levenbergmarquardt.jl:318-347) of an m×n Jacobian (n ≤ 4, column-major).
int orc_lm_damping_diagonal(int n, int mm, const double* J, int diagonal, double* out) {
    gs_method g;
    std::memset(&g, 0, sizeof(g));
    g.lm_mu0 = 1e-6;
    g.lm_scaling = diagonal ? GS_LM_SCALING_DIAGONAL : GS_LM_SCALING_NONE;
    Method<double> mt = lm_method_from(g);
    Options<double> o;
    auto run = [&](auto tag) {
        constexpr int N = decltype(tag)::value;
        Linear<N> prob{mm, J, J};
        LsqSolver<N, Linear<N>> s(prob, mt, o);
        Vec<double, N> dd = s.damping_diagonal(J);
        for (int k = 0; k < N; ++k) out[k] = dd[k];
        return 0;
    };
    switch (n) {
        case 1: return run(std::integral_constant<int, 1>{});
        case 2: return run(std::integral_constant<int, 2>{});
        case 3: return run(std::integral_constant<int, 3>{});
        case 4: return run(std::integral_constant<int, 4>{});
    }
    return GS_EINVAL;
}

// exp_family over an array (for the KATs against the C library's exp).
void orc_exp_family(const double* x, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = exp_family(x[i]);
}

// givensAlgorithm (LinearAlgebra) for the rotation KATs: out = (c, s, r).
void orc_givens(double f, double g, double* out) {
    Givens G = givens_algorithm(f, g);
    out[0] = G.c;
    out[1] = G.s;
    out[2] = G.r;
}

}  // extern "C"
