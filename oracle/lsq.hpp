// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// The least-squares half of the reference's R1 solve: `solve!!` on a LeastSquaresProblem
// (src/nonlinear/solver.jl:148-189, 301-377) with Newton's LevenbergMarquardt step
// (src/nonlinear/newton.jl:1876-1971) on the stacked Givens reduction of [J | −F]
// (src/globalization/levenbergmarquardt.jl:48-98, 162-164, 318-482, 504-584), for m ≥ n residuals
// (m a runtime size) and n ≤ 8 unknowns. Jacobians are dense m×n, column-major.
//
// The plane rotation is LinearAlgebra's `givensAlgorithm` (Julia stdlib, not under the reference
// tree; called at levenbergmarquardt.jl:374, 404, 453, 592), restated from LAPACK's DLARTG scheme it
// follows: r = √(f² + g²), c = f/r, s = g/r, with power-of-two rescaling outside [safmn2, safmx2]
// and the sign convention c ≥ 0 where |f| > |g|.
#pragma once

#include <vector>

#include "oracle.hpp"

namespace orc {

// floatmin2 of LinearAlgebra's givens: 2^-537 for Float64 (0x21a0000000000000).
inline double floatmin2_f64() {
    uint64_t b = 0x21a0000000000000ULL;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

struct Givens {
    double c, s, r;
};

inline Givens givens_algorithm(double f, double g) {
    const double safmn2 = floatmin2_f64();
    const double safmx2 = 1.0 / safmn2;
    if (g == 0.0) return {1.0, 0.0, f};
    if (f == 0.0) return {0.0, 1.0, g};
    double f1 = f, g1 = g, r, cs, sn;
    double scale = jmax(std::fabs(f1), std::fabs(g1));
    if (scale >= safmx2) {
        int count = 0;
        do {
            ++count;
            f1 *= safmn2;
            g1 *= safmn2;
            scale = jmax(std::fabs(f1), std::fabs(g1));
        } while (!(scale < safmx2 || count >= 20));
        r = std::sqrt(f1 * f1 + g1 * g1);
        cs = f1 / r;
        sn = g1 / r;
        for (int i = 0; i < count; ++i) r *= safmx2;
    } else if (scale <= safmn2) {
        int count = 0;
        do {
            ++count;
            f1 *= safmx2;
            g1 *= safmx2;
            scale = jmax(std::fabs(f1), std::fabs(g1));
        } while (!(scale > safmn2));
        r = std::sqrt(f1 * f1 + g1 * g1);
        cs = f1 / r;
        sn = g1 / r;
        for (int i = 0; i < count; ++i) r *= safmn2;
    } else {
        r = std::sqrt(f1 * f1 + g1 * g1);
        cs = f1 / r;
        sn = g1 / r;
    }
    if (std::fabs(f) > std::fabs(g) && cs < 0) {
        cs = -cs;
        sn = -sn;
        r = -r;
    }
    return {cs, sn, r};
}

// A least-squares problem of the oracle: m residuals in N unknowns.
//   int m() const;  void F(const double* x, double* r);  void J(const double* x, double* Jcm /* m×N col-major */);
//   long nf;

template <int N, class P> struct LsqSolver {
    P& prob;
    const Method<double>& mt;
    const Options<double>& o;
    const int m;
    std::vector<double> Jb, Fb, Ft, Jd;

    LsqSolver(P& p, const Method<double>& mm, const Options<double>& oo)
        : prob(p), mt(mm), o(oo), m(p.m()), Jb(size_t(p.m()) * N), Fb(p.m()), Ft(p.m()), Jd(p.m()) {}

    double rnorm_m(const double* a) const {
        // rnorm (reductions.jl:41-57) over m entries
        double s = a[0] * a[0];
        for (int i = 1; i < m; ++i) s = s + a[i] * a[i];
        if (std::isfinite(s) && s >= double(m) * floatmin_<double>()) return std::sqrt(s);
        double mx = std::fabs(a[0]);
        for (int i = 1; i < m; ++i) mx = jmax(mx, std::fabs(a[i]));
        if (mx == 0.0 || !std::isfinite(mx)) return mx;
        int e = exponent_(mx);
        double t = ldexp2(a[0], -e), acc = t * t;
        for (int i = 1; i < m; ++i) {
            t = ldexp2(a[i], -e);
            acc = acc + t * t;
        }
        return ldexp2(std::sqrt(acc), e);
    }

    // grad_ratio_and_norm (newton.jl:1287-1298) for an m×N Jacobian: ‖J‖_F over the column-major
    // entries; Jᴴ F as StaticArrays' muladd fold over the m rows.
    void grad_ratio(const double* J, const double* Fx, double fnorm, double& gr, double& jn) const {
        std::vector<double> flat(J, J + size_t(m) * N);
        {
            double s = flat[0] * flat[0];
            for (size_t i = 1; i < flat.size(); ++i) s = s + flat[i] * flat[i];
            if (std::isfinite(s) && s >= double(flat.size()) * floatmin_<double>())
                jn = std::sqrt(s);
            else {
                double mx = std::fabs(flat[0]);
                for (double v : flat) mx = jmax(mx, std::fabs(v));
                if (mx == 0.0 || !std::isfinite(mx))
                    jn = mx;
                else {
                    int e = exponent_(mx);
                    double acc = 0.0;
                    bool first = true;
                    for (double v : flat) {
                        double t = ldexp2(v, -e);
                        acc = first ? t * t : acc + t * t;
                        first = false;
                    }
                    jn = ldexp2(std::sqrt(acc), e);
                }
            }
        }
        if (jn == 0.0 || fnorm == 0.0) {
            gr = 0.0;
            return;
        }
        auto gvec = [&](double sc) {
            Vec<double, N> g;
            for (int k = 0; k < N; ++k) {
                double acc = J[0 + size_t(k) * m] * (Fx[0] * sc);
                for (int i = 1; i < m; ++i) acc = std::fma(J[i + size_t(k) * m], Fx[i] * sc, acc);
                g[k] = acc;
            }
            return g;
        };
        double gnorm = rnorm(gvec(1.0));
        if (std::isfinite(gnorm) || !(std::isfinite(jn) && std::isfinite(fnorm))) {
            gr = gnorm / jn / fnorm;
            return;
        }
        int e = exponent_(fnorm);
        double gs = rnorm(gvec(ldexp2(1.0, -e)));
        gr = gs / jn / ldexp2(fnorm, -e);
    }

    // lm_qr_factor! (levenbergmarquardt.jl:359-421): [J | −F] → (R0, c0).
    void qr_factor(const double* J, const double* Fx, Mat<double, N>& R0, Vec<double, N>& c0) const {
        auto R = [&](int i, int j) -> double& { return R0[i + j * N]; };
        for (int j = 0; j < N; ++j) {
            c0[j] = -Fx[j];
            for (int i = 0; i < N; ++i) R(i, j) = J[i + size_t(j) * m];
        }
        for (int j = 0; j < N; ++j) {
            for (int i = j + 1; i < N; ++i) {
                double b = R(i, j);
                if (b == 0.0) continue;
                double a = R(j, j);
                Givens G = givens_algorithm(a, b);
                R(j, j) = G.r;
                R(i, j) = 0.0;
                for (int k = j + 1; k < N; ++k) {
                    double u = R(j, k), v = R(i, k);
                    R(j, k) = G.c * u + G.s * v;
                    R(i, k) = -G.s * u + G.c * v;
                }
                double ur = c0[j], vr = c0[i];
                c0[j] = G.c * ur + G.s * vr;
                c0[i] = -G.s * ur + G.c * vr;
            }
        }
        for (int i = N; i < m; ++i) {
            double w[N];
            for (int k = 0; k < N; ++k) w[k] = J[i + size_t(k) * m];
            double wr = -Fx[i];
            for (int k = 0; k < N; ++k) {
                double b = w[k];
                if (b == 0.0) continue;
                double a = R(k, k);
                Givens G = givens_algorithm(a, b);
                R(k, k) = G.r;
                w[k] = 0.0;
                for (int l = k + 1; l < N; ++l) {
                    double u = R(k, l), v = w[l];
                    R(k, l) = G.c * u + G.s * v;
                    w[l] = -G.s * u + G.c * v;
                }
                double ur = c0[k];
                c0[k] = G.c * ur + G.s * wr;
                wr = -G.s * ur + G.c * wr;
            }
        }
    }

    // lm_damping_diagonal! (levenbergmarquardt.jl:318-341) of the m×N Jacobian J: with scaling =
    // :diagonal the column norms (rnorm over the m rows) floored at √eps(R) of the largest (at one
    // where every column is zero); with :none, ones.
    Vec<double, N> damping_diagonal(const double* J) const {
        Vec<double, N> dd;
        if (!mt.lm_diagonal) {
            dd.fill(1.0);
            return dd;
        }
        double dmax = 0.0;
        for (int j = 0; j < N; ++j) {
            dd[j] = rnorm_m(J + size_t(j) * m);
            dmax = jmax(dmax, dd[j]);
        }
        const double floor = dmax == 0.0 ? 1.0 : std::sqrt(eps_<double>()) * dmax;
        for (int j = 0; j < N; ++j) dd[j] = jmax(dd[j], floor);
        return dd;
    }

    // lm_trial (levenbergmarquardt.jl:575-584) → lm_qr_solve! (:438-482) with the damping rows
    // √μ d_j e_jᵀ.
    bool trial(const Mat<double, N>& R0, const Vec<double, N>& c0, double mu, const Vec<double, N>& dd,
               Vec<double, N>& d) const {
        double smu = std::sqrt(mu < 4 * eps_<double>() ? 0.0 : mu);  // lm_step_damping (:162-164)
        d.fill(0.0);
        if (!std::isfinite(smu)) return false;  // lm_damping_ok (:425-431)
        for (int j = 0; j < N; ++j)
            if (!std::isfinite(smu * dd[j])) return false;
        Mat<double, N> Rm = R0;
        Vec<double, N> c = c0;
        auto R = [&](int i, int j) -> double& { return Rm[i + j * N]; };
        for (int j = 0; j < N; ++j) {
            double w[N];
            for (int k = 0; k < N; ++k) w[k] = 0.0;
            w[j] = smu * dd[j];
            double wr = 0.0;
            for (int k = j; k < N; ++k) {
                double b = w[k];
                if (b == 0.0) continue;
                double a = R(k, k);
                Givens G = givens_algorithm(a, b);
                R(k, k) = G.r;
                w[k] = 0.0;
                for (int l = k + 1; l < N; ++l) {
                    double u = R(k, l), v = w[l];
                    R(k, l) = G.c * u + G.s * v;
                    w[l] = -G.s * u + G.c * v;
                }
                double ur = c[k];
                c[k] = G.c * ur + G.s * wr;
                wr = -G.s * ur + G.c * wr;
            }
        }
        for (int i = N - 1; i >= 0; --i) {
            double s = c[i];
            for (int j = i + 1; j < N; ++j) s = s - R(i, j) * d[j];
            double rii = R(i, i);
            if (rii == 0.0 || !std::isfinite(rii)) {
                d.fill(0.0);
                return false;
            }
            d[i] = s / rii;
        }
        for (int i = 0; i < N; ++i)
            if (!std::isfinite(d[i])) return false;
        return true;
    }

    struct Diag {
        double gradratio = nan_<double>();
        double jnorm = 0.0;
    };

    // lm_newton_factor (newton.jl:1881-1892) without a factorisation on the stacked path.
    Diag factor_diag(const Vec<double, N>& x, const double* Fx, double fnorm, bool& jfinite) {
        prob.J(x.data(), Jb.data());
        Diag dg;
        grad_ratio(Jb.data(), Fx, fnorm, dg.gradratio, dg.jnorm);
        jfinite = true;
        for (double v : Jb)
            if (!std::isfinite(v)) jfinite = false;
        return dg;
    }

    // The LevenbergMarquardt newton_step!! (newton.jl:1897-1971). Updates x, Fx, mu and the diagnostics.
    StepInfo<double> step(Vec<double, N>& x, std::vector<double>& Fx, double& mu, Diag& dg) {
        const double fnorm = rnorm_m(Fx.data());
        if (!std::isfinite(mu)) return step_info<double>(NONFINITE, fnorm, 0.0, 0, false, false);
        if (fnorm == 0.0) return step_info<double>(SUCCESS, 0.0, 0.0, 0, false, false);
        const double mu_min_eff = jmax(mt.lm_mu_min, 4 * eps_<double>());
        bool jfin;
        dg = factor_diag(x, Fx.data(), fnorm, jfin);
        if (!jfin) return step_info<double>(NONFINITE, fnorm, 0.0, 0, false, false);
        Mat<double, N> R0;
        Vec<double, N> c0;
        qr_factor(Jb.data(), Fx.data(), R0, c0);
        const Vec<double, N> dd = damping_diagonal(Jb.data());
        int32_t failures = 0;
        while (true) {
            Vec<double, N> d;
            bool ok = trial(R0, c0, mu, dd, d);
            if (ok) {
                // Jδ as StaticArrays' muladd fold over the N columns of each row
                for (int i = 0; i < m; ++i) {
                    double acc = Jb[i] * d[0];
                    for (int k = 1; k < N; ++k) acc = std::fma(Jb[i + size_t(k) * m], d[k], acc);
                    Jd[i] = acc;
                }
                double rd = Fx[0] * Jd[0];
                for (int i = 1; i < m; ++i) rd = rd + Fx[i] * Jd[i];
                double n2 = Jd[0] * Jd[0];
                for (int i = 1; i < m; ++i) n2 = n2 + Jd[i] * Jd[i];
                double pred = -2 * rd - n2;
                if (pred <= 0.0) {
                    for (int i = 0; i < m; ++i) Ft[i] = 1.0 * Fx[i] + 1.0 * Jd[i];
                    double fl = rnorm_m(Ft.data());
                    pred = (fnorm - fl) * (fnorm + fl);
                }
                Vec<double, N> xp;
                for (int k = 0; k < N; ++k) xp[k] = x[k] + 1.0 * d[k];
                prob.F(xp.data(), Ft.data());
                double fnp = rnorm_m(Ft.data());
                if (std::isfinite(fnp) && pred > 0.0) {
                    double ared = (fnorm - fnp) * (fnorm + fnp);
                    double rho = ared / pred;
                    if (rho > mt.lm_rho_accept) {
                        bool limited = mu > mt.lm_mu0;
                        double mu_next = jmax(mu * mt.lm_mu_decrease, mu_min_eff);
                        double stepnorm = rnorm(d);
                        x = xp;
                        std::copy(Ft.begin(), Ft.end(), Fx.begin());
                        mu = mu_next;
                        return step_info<double>(SUCCESS, fnp, stepnorm, failures, false, limited);
                    }
                }
            }
            failures += 1;
            mu = jmax(mu * mt.lm_mu_increase, mu_min_eff);
            if (!std::isfinite(mu) || mu > mt.lm_mu_max) {
                mu = jmin(mu, floatmax_<double>());
                return step_info<double>(STALLED, fnorm, 0.0, failures, false, false);
            }
        }
    }

    // stationary_converged (solver.jl:152-158) on the diagnostics dg.
    bool stationary(const Diag& dg, const Status<double>& s) const {
        if (!(s.iterations >= o.min_iterations)) return false;
        return dg.jnorm > 0.0 && s.fnorm > 0.0 && dg.gradratio <= o.g_tol;
    }

    Status<double> finish(const Diag& dg, Status<double> s) const {
        // newton_finish!! (newton.jl:1165-1183): rank −1, nullnorm NaN, gradratio of the diagnostics
        s.f_calls = 0;
        s.jvp_calls = 0;
        s.rank = -1;
        s.nullnorm = nan_<double>();
        s.gradratio = dg.gradratio;
        s.reused = false;
        return s;
    }

    // solve!! (solver.jl:277-378) on a LeastSquaresProblem with Newton + LevenbergMarquardt.
    Status<double> solve(Vec<double, N>& x) {
        std::vector<double>& Fx = Fb;
        prob.F(x.data(), Fx.data());
        const double fnorm0 = rnorm_m(Fx.data());
        double mu = mt.lm_mu0;  // reset_rule (newton.jl:678-680)
        Diag dg;                // newton_start!!: rank −1, NaN, NaN, jnorm 0
        Status<double> status;
        status.fnorm = fnorm0;
        if (!std::isfinite(fnorm0)) return finish(dg, closing_status(status, NONFINITE, false, false));
        const double ftest0 = o.norm == MAXNORM ? rmaxnorm_m(Fx.data()) : fnorm0;
        {
            Status<double> t = status;
            t.fnorm = ftest0;
            if (residual_converged(o, t, ftest0)) return finish(dg, status);
        }
        if (o.max_iterations == 0) return finish(dg, closing_status(status, MAXITERS, false, false));
        int32_t stalls = 0;
        std::vector<double> Fk(m);
        while (true) {
            Vec<double, N> xk = x;
            std::copy(Fx.begin(), Fx.end(), Fk.begin());
            Status<double> status_k = status;
            StepInfo<double> info = step(x, Fx, mu, dg);
            Status<double> rec = record(status, info);
            Diag dg_step = dg;  // state_step: the diagnostics of the factorisation at x_k
            status = rec;
            if (info.code == SINGULAR || info.code == NONFINITE || info.code == MAXITERS) return finish(dg, status);
            if (info.code == LINESEARCH_FAILED && info.stepnorm == 0.0 && !stationary(dg, status))
                return finish(dg, status);
            Status<double> tested = rec;
            if (o.norm == MAXNORM) tested.fnorm = rmaxnorm_m(Fx.data());
            if (residual_converged(o, tested, ftest0))
                return finish(dg, closing_status(status, SUCCESS, false, false));
            bool refreshed = false;
            auto refresh = [&]() {
                bool jf;
                dg = factor_diag(x, Fx.data(), rnorm_m(Fx.data()), jf);
                refreshed = true;
            };
            if (stationary(dg, status)) {
                refresh();
                if (stationary(dg, status)) return finish(dg, closing_status(status, SUCCESS, true, false));
                // keeps_step_start (solver.jl:183-189): return x_k with the status of x_k
                if (status_k.iterations >= 1 && status.fnorm >= status_k.fnorm && stationary(dg_step, status_k)) {
                    x = xk;
                    std::copy(Fk.begin(), Fk.end(), Fx.begin());
                    Status<double> d = status_k;
                    d.code = SUCCESS;
                    d.step_failures = status.step_failures;
                    d.promoted = status.promoted;
                    d.stationary = true;
                    d.contracted = false;
                    return finish(dg_step, d);
                }
            }
            bool steptested = !(info.code == STALLED || info.limited);
            double xnorm = steptested ? rnorm(x) : nan_<double>();
            bool contracted = !info.shortened && rate_converged(o, status, status_k.rate, xnorm);
            if (contracted ||
                (steptested && (step_converged(o, status, xnorm) ||
                                (info.code == SUCCESS && info.stepnorm == 0.0 && status.iterations >= o.min_iterations) ||
                                roundoff_stop(o, status, status_k.rate)))) {
                if (!refreshed) refresh();
                if (stationary(dg, status)) return finish(dg, closing_status(status, SUCCESS, true, false));
                if (contracted) return finish(dg, closing_status(status, SUCCESS, false, true));
                return finish(dg, closing_status(status, STALLED, false, false));
            }
            if (info.code == STALLED) {
                stalls += 1;
                if (stalls >= o.max_stalls) return finish(dg, closing_status(status, STALLED, false, false));
            } else {
                stalls = 0;
            }
            if (status.iterations >= o.max_iterations) return finish(dg, closing_status(status, MAXITERS, false, false));
        }
    }

    double rmaxnorm_m(const double* a) const {
        double mx = std::fabs(a[0]);
        for (int i = 1; i < m; ++i) mx = jmax(mx, std::fabs(a[i]));
        return mx;
    }
};

}  // namespace orc
