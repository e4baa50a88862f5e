// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// The C entry points of the oracle library (oracle/_build/liboracle.so): the batched solve of the
// registered families on the host, with the ABI structs of include/gsolve.h, and the building blocks
// that the known-answer tests reach one by one. The batch loop is OpenMP `schedule(static)` over
// systems, the host analogue of `Threads.@threads :static` over R1 solves
// (scripts/r2_throughput.jl:119, benchmark/devices.jl:454-458).
#include <omp.h>

#include <cstdio>
#include <cstring>

#include "../include/gsolve.h"
#include "families.hpp"

namespace orc {

template <class R> static Options<R> options_from(const gs_options_f64& g) {
    Options<R> o;
    o.f_abstol = R(g.f_abstol);
    o.f_reltol = R(g.f_reltol);
    o.x_abstol = R(g.x_abstol);
    o.x_reltol = R(g.x_reltol);
    o.min_iterations = g.min_iterations;
    o.max_iterations = g.max_iterations;
    o.max_stalls = g.max_stalls;
    o.g_tol = R(g.g_tol);
    o.norm = g.norm;
    o.kappa = R(g.kappa);
    o.roundoff = g.roundoff != 0;
    o.rate_test = g.rate_test != 0;
    return o;
}

template <class R> static Method<R> method_from(const gs_method& g) {
    Method<R> m;
    m.factor_f32 = g.factor_precision == GS_FACTOR_F32;
    m.glob = g.globalization;
    m.static_alpha = R(g.static_alpha);
    m.bt_c1 = R(g.bt_c1);
    m.bt_p = R(g.bt_p);
    m.bt_maxiter = g.bt_maxiter;
    m.refinement = g.refinement == GS_REFINEMENT;
    m.ref_eta = R(g.ref_eta);
    m.ref_maxiter = g.ref_maxiter;
    m.regularization = R(g.regularization);
    m.best = g.best != 0;
    m.reuse = g.reuse;
    m.dl_radius = R(g.dl_radius);
    m.lm_mu0 = R(g.lm_mu0);
    m.lm_mu_min = R(g.lm_mu_min);
    m.lm_mu_max = R(g.lm_mu_max);
    m.lm_mu_decrease = R(g.lm_mu_decrease);
    m.lm_mu_increase = R(g.lm_mu_increase);
    m.lm_rho_accept = R(g.lm_rho_accept);
    m.lm_diagonal = g.lm_scaling == GS_LM_SCALING_DIAGONAL;
    return m;
}

static void write_status(gs_status_f64* out, const Status<double>& s) {
    gs_status_f64 t;
    std::memset(&t, 0, sizeof(t));
    t.code = s.code;
    t.f_calls = s.f_calls;
    t.iterations = s.iterations;
    t.fnorm = s.fnorm;
    t.stepnorm = s.stepnorm;
    t.step_failures = s.step_failures;
    t.promoted = s.promoted;
    t.stationary = s.stationary;
    t.jvp_calls = s.jvp_calls;
    t.rank = s.rank;
    t.nullnorm = s.nullnorm;
    t.gradratio = s.gradratio;
    t.rate = s.rate;
    t.contracted = s.contracted;
    t.reused = s.reused;
    *out = t;
}

static void write_status(gs_status_f32* out, const Status<float>& s) {
    gs_status_f32 t;
    std::memset(&t, 0, sizeof(t));
    t.code = s.code;
    t.f_calls = s.f_calls;
    t.iterations = s.iterations;
    t.fnorm = s.fnorm;
    t.stepnorm = s.stepnorm;
    t.step_failures = s.step_failures;
    t.promoted = s.promoted;
    t.stationary = s.stationary;
    t.jvp_calls = s.jvp_calls;
    t.rank = s.rank;
    t.nullnorm = s.nullnorm;
    t.gradratio = s.gradratio;
    t.rate = s.rate;
    t.contracted = s.contracted;
    t.reused = s.reused;
    *out = t;
}

template <int N, class R, class SS, class MakeProb>
static void run_batch(const Method<R>& m, const Options<R>& o, int64_t nsys, R* X, SS* S, uint8_t* codes,
                      const int64_t* order, int nthreads, MakeProb make) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < nsys; ++i) {
        int64_t idx = order ? order[i] - 1 : i;
        auto prob = make(idx);
        Solver<R, N, decltype(prob)> solver(prob, m, o);
        Vec<R, N> x;
        for (int k = 0; k < N; ++k) x[k] = X[idx * N + k];
        Status<R> st = solver.solve(x);
        for (int k = 0; k < N; ++k) X[idx * N + k] = x[k];
        if (S) write_status(S + idx, st);
        if (codes) codes[idx] = st.code;
    }
}

// run_batch on the problems of `make`, or, with a deflation, on their Deflated wrappers.
template <int N, class R, class SS, class MakeProb>
static void run_batch_d(const gs_deflation* d, const Method<R>& m, const Options<R>& o, int64_t nsys, R* X, SS* S,
                        uint8_t* codes, const int64_t* order, int nthreads, MakeProb make) {
    if (!d) {
        run_batch<N>(m, o, nsys, X, S, codes, order, nthreads, make);
        return;
    }
    run_batch<N>(m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
        using Inner = decltype(make(idx));
        DeflatedProb<R, N, Inner> w{make(idx)};
        w.nroots = d->nroots;
        w.power = int64_t(d->power);
        w.shift = R(d->shift);
        for (int r = 0; r < 4; ++r)
            for (int i = 0; i < N; ++i) w.roots[r][i] = r < d->nroots ? R(d->roots[r * N + i]) : R(0);
        return w;
    });
}

// A Reparametrized batch: the problems of `make` in the chart coordinates of `ch`; X holds y₀ on entry
// and receives x* = g(y*) (reparametrized.jl:860-865).
template <int N, class R, class SS, class Chart, class MakeProb>
static void run_batch_chart(const Chart& ch, const Method<R>& m, const Options<R>& o, int64_t nsys, R* X, SS* S,
                            uint8_t* codes, const int64_t* order, int nthreads, MakeProb make) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < nsys; ++i) {
        const int64_t idx = order ? order[i] - 1 : i;
        ReparamProb<N, decltype(make(idx)), Chart> prob{make(idx), ch};
        Solver<R, N, decltype(prob)> solver(prob, m, o);
        Vec<R, N> y;
        for (int k = 0; k < N; ++k) y[k] = X[idx * N + k];
        Status<R> st = solver.solve(y);
        const Vec<R, N> x = ch.g(y);
        for (int k = 0; k < N; ++k) X[idx * N + k] = x[k];
        if (S) write_status(S + idx, st);
        if (codes) codes[idx] = st.code;
    }
}

template <int N, class R, class SS, class MakeProb>
static void run_batch_d(const gs_chart* c, const Method<R>& m, const Options<R>& o, int64_t nsys, R* X, SS* S,
                        uint8_t* codes, const int64_t* order, int nthreads, MakeProb make) {
    if constexpr (N == 2 && std::is_same_v<R, double>) {
        if (c->kind == GS_CHART_POLAR) {
            run_batch_chart<N>(ChartPolar{}, m, o, nsys, X, S, codes, order, nthreads, make);
            return;
        }
    }
    ChartLinear<N, R> ch;
    for (int i = 0; i < N; ++i) {
        ch.c[i] = R(c->c[i]);
        for (int j = 0; j < N; ++j) ch.B[i][j] = R(c->B[i * N + j]);
    }
    run_batch_chart<N>(ch, m, o, nsys, X, S, codes, order, nthreads, make);
}

template <int N, class R, class SS, class MakeProb>
static void run_batch_d(std::pair<const gs_deflation*, const gs_chart*> w, const Method<R>& m, const Options<R>& o,
                        int64_t nsys, R* X, SS* S, uint8_t* codes, const int64_t* order, int nthreads, MakeProb make) {
    if (w.second) {
        run_batch_d<N>(w.second, m, o, nsys, X, S, codes, order, nthreads, make);
        return;
    }
    run_batch_d<N>(w.first, m, o, nsys, X, S, codes, order, nthreads, make);
}

// The lane-per-system families (1-5) in the working precision R, records of R (Int64 codes for 2),
// deflated where `defl` is given, reparametrized where `chart` is.
template <class R, class SS>
static int solve_small_batch(const Method<R>& m, const Options<R>& o, const gs_family* fam, int64_t nsys, R* X,
                             const void* P, int64_t p_stride, SS* S, uint8_t* codes, const int64_t* order,
                             int nthreads, const gs_deflation* defl_ = nullptr, const gs_chart* chart = nullptr) {
    const auto defl = std::make_pair(defl_, chart);
    const char* Pb = static_cast<const char*>(P);
    auto rec = [&](int64_t idx) -> const char* { return Pb + (p_stride ? idx * p_stride : 0); };
    switch (fam->id) {
        case GS_FAM_COUPLED2:
            run_batch_d<2>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const R* p = reinterpret_cast<const R*>(rec(idx));
                return Coupled2<R>{p[0], p[1]};
            });
            return 0;
        case GS_FAM_HETERO2:
            run_batch_d<2>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const int64_t* p = reinterpret_cast<const int64_t*>(rec(idx));
                return Hetero2<R>{p[0]};
            });
            return 0;
        case GS_FAM_SCALAR_SQUARE:
            run_batch_d<1>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const R* p = reinterpret_cast<const R*>(rec(idx));
                return ScalarSquare<R>{p[0]};
            });
            return 0;
        case GS_FAM_TANH_LINEAR3:
            run_batch_d<3>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const R* p = reinterpret_cast<const R*>(rec(idx));
                TanhLinear3<R> pr;
                for (int k = 0; k < 9; ++k) pr.A[k] = p[k];
                for (int k = 0; k < 3; ++k) pr.b[k] = p[9 + k];
                return pr;
            });
            return 0;
        case GS_FAM_QUAD_MAP2:
            run_batch_d<2>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const R* p = reinterpret_cast<const R*>(rec(idx));
                QuadMap2<R> pr;
                for (int k = 0; k < 4; ++k) pr.A[k] = p[k];
                for (int k = 0; k < 2; ++k) pr.b[k] = p[4 + k];
                for (int k = 0; k < 2; ++k) pr.c[k] = p[6 + k];
                return pr;
            });
            return 0;
        case GS_FAM_DENSE_SIN: {
            const R* A = static_cast<const R*>(fam->data);
            if (!A) return GS_EINVAL;
            auto run_n = [&](auto nn) {
                constexpr int NN = decltype(nn)::value;
                run_batch<NN>(m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                    const R* p = reinterpret_cast<const R*>(rec(idx));
                    DenseSin<R, NN> pr;
                    pr.A = A;
                    for (int k = 0; k < NN; ++k) pr.b[k] = p[k];
                    return pr;
                });
                return 0;
            };
            if (defl.first || defl.second) return GS_EFAMILY;
            switch (fam->n) {
                case 2: return run_n(std::integral_constant<int, 2>{});
                case 4: return run_n(std::integral_constant<int, 4>{});
                case 6: return run_n(std::integral_constant<int, 6>{});
                case 8: return run_n(std::integral_constant<int, 8>{});
                case 12: return run_n(std::integral_constant<int, 12>{});
                case 16: return run_n(std::integral_constant<int, 16>{});
                case 24: return run_n(std::integral_constant<int, 24>{});
                case 32: return run_n(std::integral_constant<int, 32>{});
                case 48: return run_n(std::integral_constant<int, 48>{});
                case 64: return run_n(std::integral_constant<int, 64>{});
            }
            return GS_EINVAL;
        }
        case GS_FAM_HENON_HEILES_MIDPOINT: {
            const R h = R(fam->shared[0]);
            run_batch_d<4>(defl, m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const R* p = reinterpret_cast<const R*>(rec(idx));
                HenonHeilesMidpoint<R> pr;
                pr.h = h;
                for (int k = 0; k < 4; ++k) pr.xk[k] = p[k];
                return pr;
            });
            return 0;
        }
    }
    return GS_EFAMILY;
}

// A StationarityProblem batch over the action of `make`: each system's row of X is its full state
// (A::M entries); the solve runs on the unknowns x[U] and writes back the full state x(u*)
// (stationarity.jl:530-540).
template <int N, class R, class SS, class MakeAct>
static void run_batch_stationarity(const Method<R>& m, const Options<R>& o, const gs_stationarity& t, int64_t nsys,
                                   R* X, SS* S, uint8_t* codes, const int64_t* order, int nthreads, MakeAct make) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < nsys; ++i) {
        const int64_t idx = order ? order[i] - 1 : i;
        using Act = decltype(make(idx));
        constexpr int M = Act::M;
        StationarityProb<R, N, Act> prob{make(idx)};
        for (int k = 0; k < M; ++k) prob.xref[k] = X[idx * M + k];
        for (int k = 0; k < N; ++k) {
            prob.e[k] = t.n_eq == 0 ? k : t.e[k] - 1;
            prob.u[k] = t.n_eq == 0 ? k : t.u[k] - 1;
        }
        prob.has_b = t.b_kind != GS_DATUM_NONE;
        if (prob.has_b) {
            if (t.b_kind == GS_DATUM_PER_SYSTEM) {
                const R* bp = reinterpret_cast<const R*>(reinterpret_cast<const char*>(t.b_data) + idx * t.b_stride);
                for (int k = 0; k < N; ++k) prob.b[k] = bp[k];
            } else {
                for (int k = 0; k < N; ++k) prob.b[k] = R(t.b_shared[k]);
            }
        }
        Vec<R, N> u = prob.extract(prob.xref);
        Solver<R, N, decltype(prob)> solver(prob, m, o);
        Status<R> st = solver.solve(u);
        const Vec<R, M> x = solver.prob.embed(u);
        for (int k = 0; k < M; ++k) X[idx * M + k] = x[k];
        if (S) write_status(S + idx, st);
        if (codes) codes[idx] = st.code;
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

// The host restatement of gs_solve_batch_f64 (include/gsolve.h) on host pointers.
int orc_solve_batch_f64(const gs_method* method, const gs_family* fam, int64_t nsys, double* X,
                        const void* P, int64_t p_stride, gs_status_f64* S, uint8_t* codes,
                        const int64_t* order, const gs_options_f64* options, int nthreads) {
    if (!method || !fam || !options || nsys < 0) return GS_EINVAL;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<double> m = method_from<double>(*method);
    Options<double> o = options_from<double>(*options);
    const char* Pb = static_cast<const char*>(P);
    auto rec = [&](int64_t idx) -> const char* { return Pb + (p_stride ? idx * p_stride : 0); };
    switch (fam->id) {
        case GS_FAM_FPU_GL3: {
            const double h = fam->shared[0], beta = fam->shared[1];
            run_batch<48>(m, o, nsys, X, S, codes, order, nthreads, [&](int64_t idx) {
                const double* p = reinterpret_cast<const double*>(rec(idx));
                FpuGl3<double> pr;
                pr.h = h;
                pr.beta = beta;
                for (int k = 0; k < 16; ++k) {
                    pr.q[k] = p[k];
                    pr.p[k] = p[16 + k];
                }
                return pr;
            });
            return 0;
        }
        default:
            return solve_small_batch<double>(m, o, fam, nsys, X, P, p_stride, S, codes, order, nthreads);
    }
}

// The host restatement of gs_solve_batch_deflated_f64 (include/gsolve.h).
int orc_solve_batch_deflated_f64(const gs_method* method, const gs_family* fam, const gs_deflation* defl,
                                 int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                 uint8_t* codes, const int64_t* order, const gs_options_f64* options, int nthreads) {
    if (!method || !fam || !options || !defl || nsys < 0 || defl->nroots < 1 || defl->nroots > 4) return GS_EINVAL;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<double> m = method_from<double>(*method);
    Options<double> o = options_from<double>(*options);
    return solve_small_batch<double>(m, o, fam, nsys, X, P, p_stride, S, codes, order, nthreads, defl);
}

// The host restatement of gs_solve_batch_stationarity_f64 (include/gsolve.h), |E| = 2 or 6.
int orc_solve_batch_stationarity_f64(const gs_method* method, const gs_family* act, const gs_stationarity* stat,
                                     int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                     uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                     int nthreads) {
    if (!method || !act || !options || !stat || nsys < 0 || act->id != GS_ACT_HENON_HEILES) return GS_EINVAL;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<double> m = method_from<double>(*method);
    Options<double> o = options_from<double>(*options);
    const char* Pb = static_cast<const char*>(P);
    auto make = [&](int64_t idx) {
        const double* p = reinterpret_cast<const double*>(Pb + (p_stride ? idx * p_stride : 0));
        return HenonHeilesAction<double>{p[0], p[1]};
    };
    const int ne = stat->n_eq == 0 ? 6 : stat->n_eq;
    if (ne == 2)
        run_batch_stationarity<2>(m, o, *stat, nsys, X, S, codes, order, nthreads, make);
    else if (ne == 6)
        run_batch_stationarity<6>(m, o, *stat, nsys, X, S, codes, order, nthreads, make);
    else
        return GS_EFAMILY;
    return 0;
}

// ∇S (6) and ∇²S (6×6, row-major) of the Hénon–Heiles discrete action at x, for the KATs.
void orc_hh_action_eval(double h, double lam, const double* x, double* g, double* H) {
    HenonHeilesAction<double> a{h, lam};
    Vec<double, 6> xv;
    for (int k = 0; k < 6; ++k) xv[k] = x[k];
    const Vec<double, 6> gv = a.grad(xv);
    double Hm[6][6];
    a.hess(xv, Hm);
    for (int i = 0; i < 6; ++i) {
        g[i] = gv[i];
        for (int j = 0; j < 6; ++j) H[6 * i + j] = Hm[i][j];
    }
}

// The host restatement of gs_solve_batch_reparametrized_f64 (include/gsolve.h).
int orc_solve_batch_reparametrized_f64(const gs_method* method, const gs_family* fam, const gs_chart* chart,
                                       int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                       uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                       int nthreads) {
    if (!method || !fam || !options || !chart || nsys < 0) return GS_EINVAL;
    if (chart->kind == GS_CHART_POLAR && fam->n != 2) return GS_EFAMILY;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<double> m = method_from<double>(*method);
    Options<double> o = options_from<double>(*options);
    return solve_small_batch<double>(m, o, fam, nsys, X, P, p_stride, S, codes, order, nthreads, nullptr, chart);
}

// Julia's Float64 x^n for an integer n (the deflation's power), for the KATs.
void orc_pow_int(const double* x, int64_t n, double* out, int64_t len) {
    for (int64_t i = 0; i < len; ++i) out[i] = pow_int(x[i], n);
}

// The host restatement of gs_solve_batch_f32 (include/gsolve.h): the same solve in Float32, the
// method's floats and the options converted to Float32 (ToReal{Float32}, Options{Float32}).
int orc_solve_batch_f32(const gs_method* method, const gs_family* fam, int64_t nsys, float* X, const void* P,
                        int64_t p_stride, gs_status_f32* S, uint8_t* codes, const int64_t* order,
                        const gs_options_f64* options, int nthreads) {
    if (!method || !fam || !options || nsys < 0) return GS_EINVAL;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<float> m = method_from<float>(*method);
    m.factor_f32 = false;  // factor_eltype(LUFactorization{Float32}, Float32) is the working type
    if (m.dl_radius != 0.0f) m.dl_radius = m.dl_radius > 1e10f ? 1e10f : (m.dl_radius < 1e-6f ? 1e-6f : m.dl_radius);
    Options<float> o = options_from<float>(*options);
    return solve_small_batch<float>(m, o, fam, nsys, X, P, p_stride, S, codes, order, nthreads);
}

// The host restatement of gs_solve_batch_stationarity_f32 (include/gsolve.h): the Float64 batch's
// solve in Float32, the action's records and a per-system datum holding floats.
int orc_solve_batch_stationarity_f32(const gs_method* method, const gs_family* act, const gs_stationarity* stat,
                                     int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                     uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                     int nthreads) {
    if (!method || !act || !options || !stat || nsys < 0 || act->id != GS_ACT_HENON_HEILES) return GS_EINVAL;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<float> m = method_from<float>(*method);
    m.factor_f32 = false;
    if (m.dl_radius != 0.0f) m.dl_radius = m.dl_radius > 1e10f ? 1e10f : (m.dl_radius < 1e-6f ? 1e-6f : m.dl_radius);
    Options<float> o = options_from<float>(*options);
    const char* Pb = static_cast<const char*>(P);
    auto make = [&](int64_t idx) {
        const float* p = reinterpret_cast<const float*>(Pb + (p_stride ? idx * p_stride : 0));
        return HenonHeilesAction<float>{p[0], p[1]};
    };
    const int ne = stat->n_eq == 0 ? 6 : stat->n_eq;
    if (ne == 2)
        run_batch_stationarity<2>(m, o, *stat, nsys, X, S, codes, order, nthreads, make);
    else if (ne == 6)
        run_batch_stationarity<6>(m, o, *stat, nsys, X, S, codes, order, nthreads, make);
    else
        return GS_EFAMILY;
    return 0;
}

// The host restatement of gs_solve_batch_reparametrized_f32 (include/gsolve.h): the linear chart in
// Float32.
int orc_solve_batch_reparametrized_f32(const gs_method* method, const gs_family* fam, const gs_chart* chart,
                                       int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                       uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                       int nthreads) {
    if (!method || !fam || !options || !chart || nsys < 0) return GS_EINVAL;
    if (chart->kind != GS_CHART_LINEAR) return GS_EFAMILY;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    Method<float> m = method_from<float>(*method);
    m.factor_f32 = false;
    if (m.dl_radius != 0.0f) m.dl_radius = m.dl_radius > 1e10f ? 1e10f : (m.dl_radius < 1e-6f ? 1e-6f : m.dl_radius);
    Options<float> o = options_from<float>(*options);
    return solve_small_batch<float>(m, o, fam, nsys, X, P, p_stride, S, codes, order, nthreads, nullptr, chart);
}

// F and J of the FPU-β Gauss–Legendre stage system at z (48), record qp = (q, p) (32): r (48) and
// Jc (48×48 column-major), for the family KATs.
void orc_fpu_gl3_eval(double h, double beta, const double* qp, const double* z, double* r, double* Jc) {
    FpuGl3<double> pr;
    pr.h = h;
    pr.beta = beta;
    for (int k = 0; k < 16; ++k) {
        pr.q[k] = qp[k];
        pr.p[k] = qp[16 + k];
    }
    Vec<double, 48> zz;
    for (int k = 0; k < 48; ++k) zz[k] = z[k];
    Vec<double, 48> F = pr.F(zz);
    Mat<double, 48> J = pr.J(zz);
    for (int k = 0; k < 48; ++k) r[k] = F[k];
    for (int k = 0; k < 48 * 48; ++k) Jc[k] = J[k];
}

// --- building blocks for the known-answer tests -------------------------------------------------

// factorize!! StaticLU on an n×n column-major matrix (n ≤ 8), in Float64 (tf32 = 0) or Float32 factors.
int orc_lu_factor(int n, const double* J, int tf32, double* factors, int32_t* ipiv, int32_t* info,
                  int32_t* scale_exponent) {
#define ORC_LU_CASE(NN)                                                                  \
    case NN: {                                                                           \
        Mat<double, NN> A;                                                               \
        for (int k = 0; k < NN * NN; ++k) A[k] = J[k];                                   \
        if (tf32) {                                                                      \
            auto F = lu_factorize<float, double, NN>(A);                                 \
            for (int k = 0; k < NN * NN; ++k) factors[k] = F.f[k];                       \
            for (int k = 0; k < NN; ++k) ipiv[k] = F.ipiv[k];                            \
            *info = F.info;                                                              \
            *scale_exponent = F.scale_exponent;                                          \
        } else {                                                                         \
            auto F = lu_factorize<double, double, NN>(A);                                \
            for (int k = 0; k < NN * NN; ++k) factors[k] = F.f[k];                       \
            for (int k = 0; k < NN; ++k) ipiv[k] = F.ipiv[k];                            \
            *info = F.info;                                                              \
            *scale_exponent = F.scale_exponent;                                          \
        }                                                                                \
        return 0;                                                                        \
    }
    switch (n) {
        ORC_LU_CASE(1)
        ORC_LU_CASE(2)
        ORC_LU_CASE(3)
        ORC_LU_CASE(4)
        ORC_LU_CASE(5)
        ORC_LU_CASE(6)
        ORC_LU_CASE(7)
        ORC_LU_CASE(8)
        default:
            return GS_EINVAL;
    }
#undef ORC_LU_CASE
}

// The LU solve J \ r through factorize!! and ldiv!! (same-type or mixed), n ≤ 8.
int orc_lu_solve(int n, const double* J, const double* r, int tf32, double* x) {
#define ORC_SOLVE_CASE(NN)                                                               \
    case NN: {                                                                           \
        Mat<double, NN> A;                                                               \
        Vec<double, NN> b;                                                               \
        for (int k = 0; k < NN * NN; ++k) A[k] = J[k];                                   \
        for (int k = 0; k < NN; ++k) b[k] = r[k];                                        \
        Vec<double, NN> y;                                                               \
        if (tf32)                                                                        \
            y = ldiv<float, double, NN>(lu_factorize<float, double, NN>(A), b);          \
        else                                                                             \
            y = ldiv<double, double, NN>(lu_factorize<double, double, NN>(A), b);        \
        for (int k = 0; k < NN; ++k) x[k] = y[k];                                        \
        return 0;                                                                        \
    }
    switch (n) {
        ORC_SOLVE_CASE(1)
        ORC_SOLVE_CASE(2)
        ORC_SOLVE_CASE(3)
        ORC_SOLVE_CASE(4)
        ORC_SOLVE_CASE(5)
        ORC_SOLVE_CASE(6)
        ORC_SOLVE_CASE(7)
        ORC_SOLVE_CASE(8)
        default:
            return GS_EINVAL;
    }
#undef ORC_SOLVE_CASE
}

// rnorm over a vector of length L (≤ 64), for the reduction KATs.
double orc_rnorm(const double* a, int L) {
#define ORC_RN(LL) \
    case LL: {     \
        std::array<double, LL> v; for (int k = 0; k < LL; ++k) v[k] = a[k]; return rnorm(v); }
    switch (L) {
        ORC_RN(1) ORC_RN(2) ORC_RN(3) ORC_RN(4) ORC_RN(8) ORC_RN(16)
        default: return nan_<double>();
    }
#undef ORC_RN
}

// rhs_exponent and matrix_exponent of a length-4 vector, for the scaling KATs.
int orc_rhs_exponent4(const double* r, int e) {
    std::array<double, 4> v{r[0], r[1], r[2], r[3]};
    return rhs_exponent(v, e);
}
int orc_headroom(int n) { return headroom(n); }

// backtrack_step (backtracking.jl:320-340), for the interpolation KATs.
double orc_backtrack_step(double phi0, double d0, double a, double pa, double ap, double pp, double p) {
    return backtrack_step<double>(phi0, d0, a, pa, ap, pp, p);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

namespace orc {

// The fixture `fixture` (families.hpp, 101-110) in the working precision R with parameters p, passed
// to `run` as a problem; GS_EINVAL for an unknown fixture or one with no form in R.
template <class R, class Run> static int with_fixture(int fixture, const double* p, Run&& run) {
    switch (fixture) {
        case 101: return run(DiagSquare2<R>{R(p[0]), R(p[1])});
        case 102: return run(Affine2<R>{R(p[0]), R(p[1])});
        case 103: return run(InfJacobian1<R>{});
        case 104:
            if constexpr (std::is_same<R, double>::value) return run(Floor2{});
            return GS_EINVAL;
        case 105: {
            Cubic2<R> c;
            for (int k = 0; k < 4; ++k) c.A[k] = R(p[k]);
            for (int k = 0; k < 2; ++k) c.c[k] = R(p[4 + k]);
            for (int k = 0; k < 2; ++k) c.b[k] = R(p[6 + k]);
            return run(c);
        }
        case 106: return run(LinearConstJ1<R>{R(p[0]), R(p[1])});
        case 107: return run(JumpJacobian1<R>{});
        case 108: return run(FloorResidual1<R>{R(p[0])});
        case 109: return run(CycleStep1<R>{p[0] != 0.0});
        case 110: return run(JumpStep1<R>{});
        case 9: {
            QuadMap2<R> q;
            for (int k = 0; k < 4; ++k) q.A[k] = R(p[k]);
            for (int k = 0; k < 2; ++k) q.b[k] = R(p[4 + k]);
            for (int k = 0; k < 2; ++k) q.c[k] = R(p[6 + k]);
            return run(q);
        }
    }
    return GS_EINVAL;
}

template <class R> static void write_status_any(gs_status_f64* out, const Status<R>& s) {
    Status<double> d;
    d.code = s.code;
    d.f_calls = s.f_calls;
    d.iterations = s.iterations;
    d.fnorm = double(s.fnorm);
    d.stepnorm = double(s.stepnorm);
    d.step_failures = s.step_failures;
    d.promoted = s.promoted;
    d.stationary = s.stationary;
    d.jvp_calls = s.jvp_calls;
    d.rank = s.rank;
    d.nullnorm = double(s.nullnorm);
    d.gradratio = double(s.gradratio);
    d.rate = double(s.rate);
    d.contracted = s.contracted;
    d.reused = s.reused;
    write_status(out, d);
}

template <class R>
static int solve_fixture_t(int fixture, const gs_method* method, const gs_options_f64* options, double* x,
                           const double* p, gs_status_f64* S, long* counts) {
    Method<R> m = method_from<R>(*method);
    Options<R> o = options_from<R>(*options);
    return with_fixture<R>(fixture, p, [&](auto prob) {
        constexpr int N = decltype(prob)::N;
        Solver<R, N, decltype(prob)> solver(prob, m, o);
        Vec<R, N> v;
        for (int k = 0; k < N; ++k) v[k] = R(x[k]);
        Status<R> st = solver.solve(v);
        for (int k = 0; k < N; ++k) x[k] = double(v[k]);
        write_status_any<R>(S, st);
        if (counts) {
            counts[0] = solver.prob.nf;
            counts[1] = solver.prob.nj;
        }
        return 0;
    });
}

}  // namespace orc

extern "C" {

// One solve of a fixture (families.hpp, 101-110; 9 is the quadratic-map family) with parameters p; x
// is overwritten. `f32` runs it in Float32 (x, p and the status fields pass through as doubles).
// counts (may be NULL) receives the residual and the Jacobian calls of the solve.
int orc_solve_fixture_t(int fixture, int f32, const gs_method* method, const gs_options_f64* options, double* x,
                        const double* p, gs_status_f64* S, long* counts) {
    return f32 ? solve_fixture_t<float>(fixture, method, options, x, p, S, counts)
               : solve_fixture_t<double>(fixture, method, options, x, p, S, counts);
}

int orc_solve_fixture_counts(int fixture, const gs_method* method, const gs_options_f64* options, double* x,
                             const double* p, gs_status_f64* S, long* counts) {
    return solve_fixture_t<double>(fixture, method, options, x, p, S, counts);
}

// F(x) and J(x) (column-major) of a fixture in Float64, or in Float32 (f32 = 1: x and p are rounded
// to Float32, and F and J pass back as doubles).
int orc_fixture_eval(int fixture, int f32, const double* p, const double* x, double* F, double* J) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        return with_fixture<R>(fixture, p, [&](auto prob) {
            constexpr int N = decltype(prob)::N;
            Vec<R, N> v;
            for (int k = 0; k < N; ++k) v[k] = R(x[k]);
            const Vec<R, N> r = prob.F(v);
            const Mat<R, N> Jm = prob.J(v);
            for (int k = 0; k < N; ++k) F[k] = double(r[k]);
            for (int k = 0; k < N * N; ++k) J[k] = double(Jm[k]);
            return 0;
        });
    };
    return f32 ? run(0.0f) : run(0.0);
}

// One step!! (newton.jl:1671-1675) of `method` from x on a fixture, after start!! (newton.jl:1135-1144),
// as the reference's per-step tests take it (test/nonlinear/refinement.jl:406-416, 634-645): x is
// overwritten with x₊; out = {code, failures, state promoted, info promoted, fnorm, stepnorm, radius,
// residual calls (F(x) included)}.
int orc_fixture_step(int fixture, const gs_method* method, double* x, const double* p, double* out) {
    Method<double> m = method_from<double>(*method);
    Options<double> o;
    return with_fixture<double>(fixture, p, [&](auto prob) {
        constexpr int N = decltype(prob)::N;
        Solver<double, N, decltype(prob)> solver(prob, m, o);
        Vec<double, N> v;
        for (int k = 0; k < N; ++k) v[k] = x[k];
        Vec<double, N> Fx = solver.prob.F(v);
        NState<double, N> st = solver.start(v, Fx, rnorm(Fx));
        StepInfo<double> info = solver.step(v, Fx, st);
        for (int k = 0; k < N; ++k) x[k] = v[k];
        out[0] = info.code;
        out[1] = info.failures;
        out[2] = st.promoted ? 1.0 : 0.0;
        out[3] = info.promoted ? 1.0 : 0.0;
        out[4] = info.fnorm;
        out[5] = info.stepnorm;
        out[6] = st.radius;
        out[7] = double(solver.prob.nf);
        return 0;
    });
}

// refine!! (refinement.jl:180-244) on a fixture at x, as the reference's unit checks call it
// (test/nonlinear/refinement.jl:328-404): the direction d (overwritten with the result), the factor of
// kind fac_kind (0: J(x) factorised in Float32; 1: J(x) in Float64; 2: the Float32 matrix −I,
// factorised in its own precision), Refinement(eta, maxiter) or, with refine = 0, NoRefinement();
// Fx = F(x) (zero_fx = 0) or the zero vector; fnorm given (has_fnorm = 1) or formed from Fx. The
// product is J(x) v, the JVP of an analytic Jacobian. *failed receives the flag.
int orc_refine_fixture(int fixture, const double* p, const double* x, double* d, int fac_kind, int refine, double eta,
                       int maxiter, int zero_fx, int has_fnorm, double fnorm, int* failed) {
    Method<double> m;
    m.refinement = refine != 0;
    m.ref_eta = eta;
    m.ref_maxiter = maxiter;
    m.regularization = 0.0;
    Options<double> o;
    return with_fixture<double>(fixture, p, [&](auto prob) {
        constexpr int N = decltype(prob)::N;
        Solver<double, N, decltype(prob)> solver(prob, m, o);
        Vec<double, N> v, dv;
        for (int k = 0; k < N; ++k) v[k] = x[k];
        for (int k = 0; k < N; ++k) dv[k] = d[k];
        Vec<double, N> Fx = solver.prob.F(v);
        if (zero_fx)
            for (int k = 0; k < N; ++k) Fx[k] = 0.0;
        const Mat<double, N> J = solver.prob.J(v);
        const double* fn = has_fnorm ? &fnorm : nullptr;
        bool fl = false;
        if (fac_kind == 0) {
            dv = solver.template refine_with<float>(dv, lu_factorize<float, double, N>(J), nullptr, v, Fx, fl, fn);
        } else if (fac_kind == 1) {
            dv = solver.template refine_with<double>(dv, lu_factorize<double, double, N>(J), nullptr, v, Fx, fl, fn);
        } else {
            Mat<float, N> B{};
            for (int k = 0; k < N; ++k) B[k + N * k] = -1.0f;
            dv = solver.template refine_with<float>(dv, lu_factorize<float, float, N>(B), nullptr, v, Fx, fl, fn);
        }
        for (int k = 0; k < N; ++k) d[k] = dv[k];
        *failed = fl ? 1 : 0;
        return 0;
    });
}

// read_mixed_stall (newton.jl:1437-1451) on a state (promoted, slow_steps) and a step
// (fnorm, stepnorm, failures, promoted = false): out = {state promoted, slow_steps, info failures,
// info promoted, info fnorm, info stepnorm}.
void orc_read_mixed_stall(int promoted, int slow_steps, double fnorm, double stepnorm, int failures, double fnorm_old,
                          double lin_rel, double* out) {
    Method<double> m;
    Options<double> o;
    Affine2<double> prob{0.0, 0.0};
    Solver<double, 2, Affine2<double>> solver(prob, m, o);
    NState<double, 2> st;
    st.promoted = promoted != 0;
    st.slow_steps = slow_steps;
    StepInfo<double> info = step_info<double>(SUCCESS, fnorm, stepnorm, failures, false, false);
    solver.read_mixed_stall(st, info, fnorm_old, lin_rel);
    out[0] = st.promoted ? 1.0 : 0.0;
    out[1] = st.slow_steps;
    out[2] = info.failures;
    out[3] = info.promoted ? 1.0 : 0.0;
    out[4] = info.fnorm;
    out[5] = info.stepnorm;
}

// F(x) and J(x) (column-major) of family 10 at n = 16, 24, 32, 48 or 64: A column-major, b of one system.
int orc_dense_eval(int n, const double* A, const double* b, const double* x, double* F, double* J) {
    auto run = [&](auto nn) {
        constexpr int NN = decltype(nn)::value;
        DenseSin<double, NN> p;
        p.A = A;
        Vec<double, NN> v;
        for (int k = 0; k < NN; ++k) {
            p.b[k] = b[k];
            v[k] = x[k];
        }
        const Vec<double, NN> r = p.F(v);
        const Mat<double, NN> Jm = p.J(v);
        for (int k = 0; k < NN; ++k) F[k] = r[k];
        for (int k = 0; k < NN * NN; ++k) J[k] = Jm[k];
        return 0;
    };
    switch (n) {
        case 2: return run(std::integral_constant<int, 2>{});
        case 4: return run(std::integral_constant<int, 4>{});
        case 6: return run(std::integral_constant<int, 6>{});
        case 8: return run(std::integral_constant<int, 8>{});
        case 12: return run(std::integral_constant<int, 12>{});
        case 16: return run(std::integral_constant<int, 16>{});
        case 24: return run(std::integral_constant<int, 24>{});
        case 32: return run(std::integral_constant<int, 32>{});
        case 48: return run(std::integral_constant<int, 48>{});
        case 64: return run(std::integral_constant<int, 64>{});
    }
    return GS_EINVAL;
}

// The sine and cosine of family 10 (families.hpp fsincos), elementwise, in Float64 or Float32.
void orc_fsincos(const double* x, double* s, double* c, int64_t len, int f32) {
    for (int64_t i = 0; i < len; ++i) {
        if (f32) {
            float a, b;
            fsincos<float>(float(x[i]), a, b);
            s[i] = a;
            c[i] = b;
        } else {
            fsincos<double>(x[i], s[i], c[i]);
        }
    }
}

// rmaxnorm (reductions.jl:67) of a length-L vector (L ≤ 4), in Float64 or Float32 (f32 = 1: the
// entries are rounded to Float32 first).
double orc_rmaxnorm(const double* a, int L, int f32) {
    auto run = [&](auto zero) -> double {
        using R = decltype(zero);
        switch (L) {
            case 0: return double(rmaxnorm(std::array<R, 0>{}));
#define ORC_RM(LL) \
    case LL: {     \
        std::array<R, LL> v; for (int k = 0; k < LL; ++k) v[k] = R(a[k]); return double(rmaxnorm(v)); }
            ORC_RM(1) ORC_RM(2) ORC_RM(3) ORC_RM(4)
#undef ORC_RM
        }
        return nan_<double>();
    };
    return f32 ? run(0.0f) : run(0.0);
}

// One DogLeg step!! (newton.jl:1725-1737) from x0 on a DogLegPw residual (families.hpp), as the
// reference's one_step (test/globalization/dogleg.jl:60-66): out = {x₊, radius after the step, code,
// failures, stepnorm, residual calls (F(x₀) included), limited}.
int orc_dogleg_pw_step(int kind, const gs_method* method, double p0, double p1, double x0, double* out) {
    Method<double> m = method_from<double>(*method);
    Options<double> o;
    DogLegPw<double> prob{kind, p0, p1};
    Solver<double, 1, DogLegPw<double>> solver(prob, m, o);
    Vec<double, 1> x{x0};
    Vec<double, 1> Fx = prob.F(x);
    NState<double, 1> st = solver.start(x, Fx, rnorm(Fx));
    StepInfo<double> info = solver.step(x, Fx, st);
    out[0] = x[0];
    out[1] = st.radius;
    out[2] = info.code;
    out[3] = info.failures;
    out[4] = info.stepnorm;
    out[5] = double(prob.nf);
    out[6] = info.limited ? 1.0 : 0.0;
    return 0;
}

int orc_solve_fixture(int fixture, const gs_method* method, const gs_options_f64* options, double* x,
                      const double* p, gs_status_f64* S) {
    return orc_solve_fixture_counts(fixture, method, options, x, p, S, nullptr);
}

}  // extern "C"
