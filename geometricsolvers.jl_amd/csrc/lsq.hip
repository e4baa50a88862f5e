// libgsolve_hip, least-squares part: gs_solve_expsum_lsq_f64 — Newton + LevenbergMarquardt on a batch
// of exponential-sum fits of uneven sizes (K ∈ 1..4 terms, n = 2K unknowns, m ≥ n points each).
//
// The reference has no R2 form of a LeastSquaresProblem; per fit this is its R1
// `solve(LeastSquaresProblem(F, J), x₀, Newton(; globalization = LevenbergMarquardt()), p)`
// (src/nonlinear/solver.jl:148-189, 277-378; src/nonlinear/newton.jl:1876-1971;
// src/globalization/levenbergmarquardt.jl:48-98, 162-164, 359-482, 575-584).
//
// Launch structure (one call): the fits are bucketed by n and ordered by m on the device (a counting
// sort over (K, m) bins), then one persistent kernel per bucket, compiled for that n, runs one fit
// per lane; a lane whose fit ends takes the next fit of its bucket from an atomic queue at the top of
// the next LM outer step. The buckets' ranges stay on the device, so the whole sequence does not
// depend on the data: it is captured once into a hipGraph and replayed while the call's pointers,
// sizes, parameters and shapes stay the same, with one host synchronisation at the end. The shape of
// each bucket (workgroup size, resident workgroups per CU) is the package's autotuner's choice
// (set through gs_lsq_set_shape) or the default (64 threads, the kernel's occupancy).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsolve.h"
#include "lsq_solver.hpp"

namespace gs {

int set_error(int code, const std::string& msg);
// One call at a time per device, held through the synchronisation (defined with gs_solve_batch_f64).
std::unique_lock<std::recursive_mutex> device_call_lock();

namespace {

constexpr int kMaxM = 4096;  // largest fit the device form takes (a bound, not a storage size)
constexpr int kBins = 4 * (kMaxM + 1);

// ---- the per-fit solve, as a step machine --------------------------------------------------------

// A lane's slot in the workspace: the triangular factor R (upper triangle, row by row) and its
// right-hand side c of the current step, and the step's start point x_k. Entry e of lane l sits at
// base[e * stride + l], so the 64 lanes of a wave touch consecutive words. Keeping these in memory
// rather than in VGPRs lets the damped trials run without two live copies of R.
template <int N> struct Slot {
    static constexpr int kR = 0, kC = N * (N + 1) / 2, kX = kC + N, kSize = kX + N;
    double* p;
    int64_t stride;
    GS_DEV double& operator[](int e) const { return p[e * stride]; }
    static constexpr int r(int i, int j) { return i * N - (i * (i - 1)) / 2 + (j - i); }
};

// A compiler-only fence: values stored to a slot are read back from memory, not kept in registers.
GS_DEV void slot_fence() { __asm__ volatile("" ::: "memory"); }

template <int N> struct Fit {
    static constexpr int K = N / 2;
    ExpSumRow<N> row;
    Slot<N> slot;
    int64_t idx;
    double x[N];
    double fnorm, ftest0, mu;
    int32_t it, failures, stalls;
    double stepnorm, rate;
    double gr, jn;  // diagnostics of the last factorisation (NaN, 0 before any)
    // the step's start status and its factorisation's diagnostics (x_k itself sits in the slot)
    int32_t it_k;
    double fnorm_k, stepnorm_k, rate_k;
    uint8_t code;
    bool stationary, contracted;
    // the diagnostics: measured at x_k (the slot) by the last step, or at x by a refresh; with the
    // reference's ‖J‖_F fold (exact) or with column sums
    bool has_dg, dg_at_xk, dg_exact;
};

// rnorm over the m residuals of `x` (reductions.jl:41-57): one streamed pass, a second with the
// scale of the largest entry where the sum overflows or underflows.
template <int N> GS_DEV double residual_norm(const ExpSumRow<N>& row, const double (&x)[N], double* maxabs) {
    double s = 0.0, mx = 0.0;
    typename ExpSumRow<N>::Ahead ya(row, 0);
    for (int i = 0; i < row.m; ++i) {
        double f = row.F(x, i, ya.take(i));
        s = i == 0 ? f * f : s + f * f;
        mx = i == 0 ? __builtin_fabs(f) : jmax(mx, __builtin_fabs(f));
    }
    if (maxabs) *maxabs = mx;
    if (__builtin_expect(isfin(s) && s >= double(row.m) * kDblMin, 1)) return __builtin_sqrt(s);
    if (mx == 0.0 || !isfin(mx)) return mx;
    int e = exponent_of(mx);
    double t = 0.0;
    for (int i = 0; i < row.m; ++i) {
        double f = ldexp2(row.F(x, i), -e);
        t = i == 0 ? f * f : t + f * f;
    }
    return ldexp2(__builtin_sqrt(t), e);
}

// ‖J‖_F at x as the reference folds it (grad_ratio_and_norm, newton.jl:1287-1298, through rnorm,
// reductions.jl:41-57): a left fold over the entries in column-major order, one pass over the rows
// per column, and the scaled second pass where the sum overflows or underflows.
template <int N> GS_DEV double jnorm_exact(const ExpSumRow<N>& row, const double (&x)[N]) {
    double s = 0.0, mx = 0.0;
    static_for<0, N>([&](auto kk) {
        constexpr int k = kk;
        for (int i = 0; i < row.m; ++i) {
            const double v = row.J(x, i, k);
            const bool first = k == 0 && i == 0;
            s = first ? v * v : s + v * v;
            mx = first ? __builtin_fabs(v) : jmax(mx, __builtin_fabs(v));
        }
    });
    if (isfin(s) && s >= double(row.m * N) * kDblMin) return __builtin_sqrt(s);
    if (mx == 0.0 || !isfin(mx)) return mx;
    const int e = exponent_of(mx);
    double t = 0.0;
    static_for<0, N>([&](auto kk) {
        constexpr int k = kk;
        for (int i = 0; i < row.m; ++i) {
            const double v = ldexp2(row.J(x, i, k), -e);
            t = (k == 0 && i == 0) ? v * v : t + v * v;
        }
    });
    return ldexp2(__builtin_sqrt(t), e);
}

// ‖J‖_F from the column sums of J∘J that a row pass folds (each column in the reference's row order;
// the columns then added in order): the reference's value up to the order of one sum, so within
// m·n·eps relative. Where that sum overflows or underflows, the reference's fold itself.
template <int N>
GS_DEV double jnorm_from_columns(const ExpSumRow<N>& row, const double (&x)[N], const double (&col)[N]) {
    double s = col[0];
    static_for<1, N>([&](auto kk) { s = s + col[int(kk)]; });
    if (__builtin_expect(isfin(s) && s >= double(row.m * N) * kDblMin, 1)) return __builtin_sqrt(s);
    return jnorm_exact<N>(row, x);
}

// ‖Jᴴ F‖ / (‖J‖_F ‖F‖) (grad_ratio_and_norm, newton.jl:1287-1298) from Jᴴ F (the muladd fold over the
// rows, the reference's order), ‖J‖_F and ‖F‖ at x.
template <int N>
GS_DEV double grad_ratio_from(const ExpSumRow<N>& row, const double (&x)[N], const double (&g)[N], double jn,
                              double fnorm) {
    if (jn == 0.0 || fnorm == 0.0) return 0.0;
    double gg[N];
    static_for<0, N>([&](auto kk) { gg[int(kk)] = g[int(kk)]; });
    const double gn = rnorm<N>(gg);
    if (isfin(gn) || !(isfin(jn) && isfin(fnorm))) return gn / jn / fnorm;
    // Jᴴ (F 2^-e) for the overflowing product
    const int e = exponent_of(fnorm);
    const double sc = ldexp2(1.0, -e);
    for (int i = 0; i < row.m; ++i) {
        double Ji[N];
        const double Fi = row.FJ(x, i, Ji) * sc;
        static_for<0, N>([&](auto kk) {
            constexpr int k = kk;
            gg[k] = i == 0 ? Ji[k] * Fi : __builtin_fma(Ji[k], Fi, gg[k]);
        });
    }
    return rnorm<N>(gg) / jn / ldexp2(fnorm, -e);
}

// Jᴴ F and the column sums of J∘J at x, folded over the rows.
template <int N> GS_DEV void diag_rows(const ExpSumRow<N>& row, const double (&x)[N], double (&g)[N], double (&col)[N]) {
    typename ExpSumRow<N>::Ahead ya(row, 0);
    for (int i = 0; i < row.m; ++i) {
        double Ji[N];
        const double Fi = row.FJ(x, i, ya.take(i), Ji);
        static_for<0, N>([&](auto kk) {
            constexpr int k = kk;
            g[k] = i == 0 ? Ji[k] * Fi : __builtin_fma(Ji[k], Fi, g[k]);
            col[k] = i == 0 ? Ji[k] * Ji[k] : col[k] + Ji[k] * Ji[k];
        });
    }
}

// The diagnostics (gradratio, ‖J‖_F) at x with the reference's ‖J‖_F fold.
template <int N>
GS_DEV void diag_exact(const ExpSumRow<N>& row, const double (&x)[N], double fnorm, double& gr, double& jn) {
    double g[N], col[N];
    diag_rows<N>(row, x, g, col);
    jn = jnorm_exact<N>(row, x);
    gr = grad_ratio_from<N>(row, x, g, jn, fnorm);
}

// refresh_diagnostics!! (newton.jl:771-783) for LevenbergMarquardt: the diagnostics re-measured at x.
template <int N> GS_DEV void diag_pass(Fit<N>& f) {
    double g[N], col[N];
    diag_rows<N>(f.row, f.x, g, col);
    f.jn = jnorm_from_columns<N>(f.row, f.x, col);
    f.gr = grad_ratio_from<N>(f.row, f.x, g, f.jn, f.fnorm);
    f.has_dg = true;
    f.dg_at_xk = false;
    f.dg_exact = false;
}

struct StepOut {
    uint8_t code;
    double fnorm, stepnorm;
    int32_t failures;
    bool limited;
};

// Pass B of a trial (newton.jl:1937-1950): one stream over the rows for ⟨F, Jδ⟩ and ‖Jδ‖² (left
// folds, Jδ as the muladd fold over the columns) and the trial residual F(x + δ) with its sum of
// squares and largest entry. With DIAG it also folds Jᴴ F and the column sums of J∘J for the
// diagnostics of the step's factorisation, whose rows it evaluates at x anyway.
struct PassB {
    double rd, n2, fp2, fpmax;
};
template <bool DIAG, int N>
GS_DEV PassB pass_b(const ExpSumRow<N>& row, const double (&x)[N], const double (&d)[N], const double (&xp)[N],
                    double (&g)[N], double (&col)[N]) {
    PassB o{0.0, 0.0, 0.0, 0.0};
    typename ExpSumRow<N>::Ahead ya(row, 0);
    for (int i = 0; i < row.m; ++i) {
        double Ji[N];
        const double yi = ya.take(i);
        const double Fi = row.FJ(x, i, yi, Ji);
        double jd = Ji[0] * d[0];
        static_for<1, N>([&](auto kk) { jd = __builtin_fma(Ji[int(kk)], d[int(kk)], jd); });
        const double fp = row.F(xp, i, yi);
        const bool first = i == 0;
        o.rd = first ? Fi * jd : o.rd + Fi * jd;
        o.n2 = first ? jd * jd : o.n2 + jd * jd;
        o.fp2 = first ? fp * fp : o.fp2 + fp * fp;
        o.fpmax = first ? __builtin_fabs(fp) : jmax(o.fpmax, __builtin_fabs(fp));
        if constexpr (DIAG) {
            static_for<0, N>([&](auto kk) {
                constexpr int k = kk;
                g[k] = first ? Ji[k] * Fi : __builtin_fma(Ji[k], Fi, g[k]);
                col[k] = first ? Ji[k] * Ji[k] : col[k] + Ji[k] * Ji[k];
            });
        }
    }
    return o;
}

// ‖F + J δ‖ (rnorm with its scaled path), for a predicted reduction that rounds to ≤ 0.
template <int N> GS_DEV double linear_norm(const ExpSumRow<N>& row, const double (&x)[N], const double (&d)[N]) {
    double s = 0.0, mx = 0.0;
    typename ExpSumRow<N>::Ahead ya(row, 0);
    for (int i = 0; i < row.m; ++i) {
        double Ji[N];
        const double Fi = row.FJ(x, i, ya.take(i), Ji);
        double jd = Ji[0] * d[0];
        static_for<1, N>([&](auto kk) { jd = __builtin_fma(Ji[int(kk)], d[int(kk)], jd); });
        const double l = 1.0 * Fi + 1.0 * jd;
        s = i == 0 ? l * l : s + l * l;
        mx = i == 0 ? __builtin_fabs(l) : jmax(mx, __builtin_fabs(l));
    }
    if (isfin(s) && s >= double(row.m) * kDblMin) return __builtin_sqrt(s);
    if (mx == 0.0 || !isfin(mx)) return mx;
    const int e = exponent_of(mx);
    double t = 0.0;
    for (int i = 0; i < row.m; ++i) {
        double Ji[N];
        const double Fi = row.FJ(x, i, Ji);
        double jd = Ji[0] * d[0];
        static_for<1, N>([&](auto kk) { jd = __builtin_fma(Ji[int(kk)], d[int(kk)], jd); });
        const double l = ldexp2(1.0 * Fi + 1.0 * jd, -e);
        t = i == 0 ? l * l : t + l * l;
    }
    return ldexp2(__builtin_sqrt(t), e);
}

// lm_damping_diagonal (levenbergmarquardt.jl:318-341) from the column sums of squares col2 (each a
// left fold over the rows, as rnorm's): Marquardt's column norms floored at √eps of the largest, or
// at one where every column is zero. A column sum that overflows or underflows takes rnorm's scaled
// second pass over that column.
template <int N>
GS_DEV void damping_diagonal(const ExpSumRow<N>& row, const double (&x)[N], const double (&col2)[N], double (&dd)[N]) {
    double dmax = 0.0;
    static_for<0, N>([&](auto kk) {
        constexpr int k = kk;
        const double s = col2[k];
        double d;
        if (__builtin_expect(isfin(s) && s >= double(row.m) * kDblMin, 1)) {
            d = __builtin_sqrt(s);
        } else {
            double mx = 0.0;
            for (int i = 0; i < row.m; ++i) {
                const double v = __builtin_fabs(row.J(x, i, k));
                mx = i == 0 ? v : jmax(mx, v);
            }
            if (mx == 0.0 || !isfin(mx)) {
                d = mx;
            } else {
                const int e = exponent_of(mx);
                double t = 0.0;
                for (int i = 0; i < row.m; ++i) {
                    const double v = ldexp2(row.J(x, i, k), -e);
                    t = i == 0 ? v * v : t + v * v;
                }
                d = ldexp2(__builtin_sqrt(t), e);
            }
        }
        dd[k] = d;
        dmax = jmax(dmax, d);
    });
    const double floor = dmax == 0.0 ? 1.0 : 1.4901161193847656e-8 * dmax;  // √eps(Float64) = 2^-26
    static_for<0, N>([&](auto kk) { dd[int(kk)] = jmax(dd[int(kk)], floor); });
}

// The LevenbergMarquardt newton_step!! (newton.jl:1897-1971) of one fit; SCALED: scaling = :diagonal.
template <int N, bool SCALED> GS_DEV StepOut lm_step(Fit<N>& f, const LsqParams& q) {
    const double fnorm = f.fnorm;
    if (!isfin(f.mu)) return StepOut{kLsqNonfinite, fnorm, 0.0, 0, false};
    if (fnorm == 0.0) return StepOut{kLsqSuccess, 0.0, 0.0, 0, false};
    // the reduction of [J | −F] (lm_qr_factor!, levenbergmarquardt.jl:359-421) into the lane's slot,
    // and with SCALED the column sums of J∘J in the rows' order
    bool jfinite = true;
    double col2[N];
    {
        double R[N][N], c[N];
        static_for<0, N>([&](auto ii) {
            constexpr int i = ii;
            double Ji[N];
            const double Fi = f.row.FJ(f.x, i, Ji);
            static_for<0, N>([&](auto kk) {
                constexpr int k = kk;
                jfinite &= isfin(Ji[k]);
                R[i][k] = Ji[k];
                if (SCALED) col2[k] = i == 0 ? Ji[k] * Ji[k] : col2[k] + Ji[k] * Ji[k];
            });
            c[i] = -Fi;
        });
        // the first n rows, column by column
        static_for<0, N>([&](auto jj) {
            constexpr int j = jj;
            static_for<j + 1, N>([&](auto ii) {
                constexpr int i = ii;
                const double b = R[i][j];
                if (__all(givens_plain_ok(R[j][j], b))) {
                    double cs, sn, r;
                    givens_plain(R[j][j], b, cs, sn, r);
                    R[j][j] = r;
                    R[i][j] = 0.0;
                    static_for<j + 1, N>([&](auto kk) {
                        constexpr int k = kk;
                        const double u = R[j][k], v = R[i][k];
                        R[j][k] = cs * u + sn * v;
                        R[i][k] = -sn * u + cs * v;
                    });
                    const double ur = c[j], vr = c[i];
                    c[j] = cs * ur + sn * vr;
                    c[i] = -sn * ur + cs * vr;
                    return;
                }
                const bool skip = b == 0.0;
                double cs, sn, r;
                givens(R[j][j], b, cs, sn, r);
                R[j][j] = skip ? R[j][j] : r;
                R[i][j] = 0.0;
                static_for<j + 1, N>([&](auto kk) {
                    constexpr int k = kk;
                    const double u = R[j][k], v = R[i][k];
                    R[j][k] = skip ? u : cs * u + sn * v;
                    R[i][k] = skip ? v : -sn * u + cs * v;
                });
                const double ur = c[j], vr = c[i];
                c[j] = skip ? ur : cs * ur + sn * vr;
                c[i] = skip ? vr : -sn * ur + cs * vr;
            });
        });
        // the remaining rows, streamed in order
        typename ExpSumRow<N>::Ahead ya(f.row, N);
#pragma unroll 1
        for (int i = N; i < f.row.m; ++i) {
            double w[N];
            const double Fi = f.row.FJ(f.x, i, ya.take(i), w);
            static_for<0, N>([&](auto kk) {
                constexpr int k = kk;
                jfinite &= isfin(w[k]);
                if (SCALED) col2[k] = col2[k] + w[k] * w[k];
            });
            double wr = -Fi;
            rotate_in<0>(R, c, w, wr);
        }
        static_for<0, N>([&](auto ii) {
            constexpr int i = ii;
            f.slot[Slot<N>::kC + i] = c[i];
            static_for<i, N>([&](auto jj) { f.slot[Slot<N>::kR + Slot<N>::r(i, int(jj))] = R[i][int(jj)]; });
        });
    }
    if (!jfinite) {
        diag_pass<N>(f);
        return StepOut{kLsqNonfinite, fnorm, 0.0, 0, false};
    }
    double dd[N];
    if (SCALED) damping_diagonal<N>(f.row, f.x, col2, dd);
    int32_t failures = 0;
    double mu = f.mu;
    bool diag_done = false;
    while (true) {
        // lm_trial (:575-584): the damping rows √μ d_j e_jᵀ (d_j = 1 without SCALED) eliminated
        // against a copy of (R, c); a damping entry that is not finite fails the trial (:425-431)
        const double smu = __builtin_sqrt(mu < 4.0 * kDblEps ? 0.0 : mu);
        bool ok = isfin(smu);
        if (SCALED) static_for<0, N>([&](auto kk) { ok &= isfin(smu * dd[int(kk)]); });
        double d[N];
        static_for<0, N>([&](auto kk) { d[int(kk)] = 0.0; });
        if (ok) {
            slot_fence();
            double Rt[N][N], ct[N];
            static_for<0, N>([&](auto aa) {
                constexpr int a = aa;
                ct[a] = f.slot[Slot<N>::kC + a];
                static_for<0, N>([&](auto bb) {
                    constexpr int b = bb;
                    Rt[a][b] = b >= a ? f.slot[Slot<N>::kR + Slot<N>::r(a, b)] : 0.0;
                });
            });
            static_for<0, N>([&](auto jj) {
                constexpr int j = jj;
                double w[N];
                static_for<0, N>([&](auto kk) { w[int(kk)] = 0.0; });
                w[j] = SCALED ? smu * dd[j] : smu * 1.0;
                double wr = 0.0;
                rotate_in<j>(Rt, ct, w, wr);
            });
            static_for<0, N>([&](auto ii) {
                constexpr int i = N - 1 - int(ii);
                double s = ct[i];
                static_for<i + 1, N>([&](auto jj) { s = s - Rt[i][int(jj)] * d[int(jj)]; });
                const double rii = Rt[i][i];
                if (rii == 0.0 || !isfin(rii)) ok = false;
                d[i] = s / rii;
            });
            if (!ok) static_for<0, N>([&](auto kk) { d[int(kk)] = 0.0; });
            static_for<0, N>([&](auto kk) { ok &= isfin(d[int(kk)]); });
        }
        if (ok) {
            // pred, the trial residual and ared, streamed over the rows
            double xp[N];
            static_for<0, N>([&](auto kk) { xp[int(kk)] = f.x[int(kk)] + 1.0 * d[int(kk)]; });
            PassB b;
            if (!diag_done) {
                double g[N], col[N];
                b = pass_b<true, N>(f.row, f.x, d, xp, g, col);
                f.jn = jnorm_from_columns<N>(f.row, f.x, col);
                f.gr = grad_ratio_from<N>(f.row, f.x, g, f.jn, fnorm);
                f.has_dg = true;
                f.dg_at_xk = true;
                f.dg_exact = false;
                diag_done = true;
            } else {
                double g[N], col[N];
                b = pass_b<false, N>(f.row, f.x, d, xp, g, col);
            }
            double pred = -2.0 * b.rd - b.n2;
            if (pred <= 0.0) {
                const double fl = linear_norm<N>(f.row, f.x, d);
                pred = (fnorm - fl) * (fnorm + fl);
            }
            double fnp;
            if (isfin(b.fp2) && b.fp2 >= double(f.row.m) * kDblMin) {
                fnp = __builtin_sqrt(b.fp2);
            } else if (b.fpmax == 0.0 || !isfin(b.fpmax)) {
                fnp = b.fpmax;
            } else {
                const int e = exponent_of(b.fpmax);
                double t = 0.0;
                for (int i = 0; i < f.row.m; ++i) {
                    const double v = ldexp2(f.row.F(xp, i), -e);
                    t = i == 0 ? v * v : t + v * v;
                }
                fnp = ldexp2(__builtin_sqrt(t), e);
            }
            if (isfin(fnp) && pred > 0.0) {
                const double ared = (fnorm - fnp) * (fnorm + fnp);
                const double rho = ared / pred;
                if (rho > q.rho_accept) {
                    const bool limited = mu > q.mu0;
                    f.mu = jmax(mu * q.mu_dec, q.mu_min_eff);
                    const double stepnorm = rnorm<N>(d);
                    static_for<0, N>([&](auto kk) { f.x[int(kk)] = xp[int(kk)]; });
                    f.fnorm = fnp;
                    return StepOut{kLsqSuccess, fnp, stepnorm, failures, limited};
                }
            }
        }
        failures += 1;
        mu = jmax(mu * q.mu_inc, q.mu_min_eff);
        if (!isfin(mu) || mu > q.mu_max) {
            f.mu = jmin(mu, 1.7976931348623157e308);
            if (!diag_done) diag_pass<N>(f);
            return StepOut{kLsqStalled, fnorm, 0.0, failures, false};
        }
    }
}

// The start of a fit's solve (solver.jl:280-298 with newton_start!!, newton.jl:1135-1144): true when
// the solve has already ended (f.code set).
template <int N> GS_DEV bool fit_start(Fit<N>& f, const LsqParams& q) {
    double mx;
    const double fnorm0 = residual_norm<N>(f.row, f.x, &mx);
    f.fnorm = fnorm0;
    f.mu = q.mu0;
    f.it = 0;
    f.failures = 0;
    f.stalls = 0;
    f.stepnorm = 0.0;
    f.rate = dnan();
    f.gr = dnan();
    f.jn = 0.0;
    f.has_dg = false;
    f.dg_at_xk = false;
    f.dg_exact = false;
    f.stationary = false;
    f.contracted = false;
    if (!isfin(fnorm0)) {
        f.code = kLsqNonfinite;
        return true;
    }
    f.ftest0 = q.norm_max ? mx : fnorm0;
    if (0 >= q.min_iterations && f.ftest0 <= jmax(q.f_abstol, q.f_reltol * f.ftest0)) {
        f.code = kLsqSuccess;
        return true;
    }
    if (q.max_iterations == 0) {
        f.code = kLsqMaxIters;
        return true;
    }
    return false;
}

// The point at which diagnostics were measured: x_k (the slot) or x.
template <int N> GS_DEV void dg_point(const Fit<N>& f, bool at_xk, double (&p)[N]) {
    if (at_xk) {
        slot_fence();
        static_for<0, N>([&](auto kk) { p[int(kk)] = f.slot[Slot<N>::kX + int(kk)]; });
    } else {
        static_for<0, N>([&](auto kk) { p[int(kk)] = f.x[int(kk)]; });
    }
}

// stationary_converged (solver.jl:152-158): iterations ≥ min_iterations, ‖J‖_F > 0, ‖F‖ > 0 (of the
// status, fn_status) and gradratio ≤ g_tol, on diagnostics (gr, jn) measured at the point `at_xk`
// with ‖F‖ = fn_dg there. Where gr lies within 1e-10 (relative) of g_tol, (gr, jn) are first
// re-measured with the reference's ‖J‖_F fold (`exact`), so that the decision is the reference's.
template <int N>
GS_DEV bool stationary_dg(const Fit<N>& f, const LsqParams& q, int32_t it, double fn_status, double& gr, double& jn,
                          bool& exact, bool at_xk, double fn_dg) {
    if (!(it >= q.min_iterations) || !(jn > 0.0 && fn_status > 0.0)) return false;
    if (!exact && gr > q.g_tol * (1.0 - 1e-10) && gr <= q.g_tol * (1.0 + 1e-10)) {
        double p[N];
        dg_point<N>(f, at_xk, p);
        diag_exact<N>(f.row, p, fn_dg, gr, jn);
        exact = true;
    }
    return gr <= q.g_tol;
}

// One pass of the solve!! loop on a LeastSquaresProblem (solver.jl:301-377): a step, its record, and
// every exit. Returns true when the fit's solve has ended.
template <int N, bool SCALED> GS_DEV bool fit_iterate(Fit<N>& f, const LsqParams& q) {
    static_for<0, N>([&](auto kk) { f.slot[Slot<N>::kX + int(kk)] = f.x[int(kk)]; });
    f.it_k = f.it;
    f.fnorm_k = f.fnorm;
    f.stepnorm_k = f.stepnorm;
    f.rate_k = f.rate;
    StepOut s = lm_step<N, SCALED>(f, q);
    // record (status.jl:245-262)
    f.rate = f.stepnorm > 0.0 ? s.stepnorm / f.stepnorm : dnan();
    f.code = s.code;
    f.it += 1;
    f.fnorm = s.fnorm;
    f.stepnorm = s.stepnorm;
    f.failures += s.failures;
    if (s.code == kLsqNonfinite) return true;
    double ftest = f.fnorm;
    if (q.norm_max) {
        double mx;
        residual_norm<N>(f.row, f.x, &mx);
        ftest = mx;
    }
    if (f.it >= q.min_iterations && ftest <= jmax(q.f_abstol, q.f_reltol * f.ftest0)) {
        f.code = kLsqSuccess;
        return true;
    }
    bool refreshed = false;
    const bool stat_step = stationary_dg<N>(f, q, f.it, f.fnorm, f.gr, f.jn, f.dg_exact, f.dg_at_xk,
                                            f.dg_at_xk ? f.fnorm_k : f.fnorm);
    // state_step: the diagnostics of the step's factorisation, at x_k
    double gr_step = f.gr, jn_step = f.jn;
    bool ex_step = f.dg_exact;
    const bool at_step = f.dg_at_xk;
    if (stat_step) {
        diag_pass<N>(f);
        refreshed = true;
        if (stationary_dg<N>(f, q, f.it, f.fnorm, f.gr, f.jn, f.dg_exact, false, f.fnorm)) {
            f.code = kLsqSuccess;
            f.stationary = true;
            return true;
        }
        // keeps_step_start (solver.jl:183-189): the step left a stationary start point without lowering ‖F‖
        if (f.it_k >= 1 && f.fnorm >= f.fnorm_k &&
            stationary_dg<N>(f, q, f.it_k, f.fnorm_k, gr_step, jn_step, ex_step, at_step, f.fnorm_k)) {
            slot_fence();
            static_for<0, N>([&](auto kk) { f.x[int(kk)] = f.slot[Slot<N>::kX + int(kk)]; });
            f.it = f.it_k;
            f.fnorm = f.fnorm_k;
            f.stepnorm = f.stepnorm_k;
            f.rate = f.rate_k;
            f.gr = gr_step;
            f.jn = jn_step;
            f.dg_exact = ex_step;
            f.dg_at_xk = true;
            f.code = kLsqSuccess;
            f.stationary = true;
            return true;
        }
    }
    const bool steptested = !(s.code == kLsqStalled || s.limited);
    const double xnorm = steptested ? rnorm<N>(f.x) : dnan();
    const double xtol = jmax(q.x_abstol, q.x_reltol * xnorm);
    const bool past_min = f.it >= q.min_iterations;
    bool contracted = false;
    if (q.rate_test && past_min) {
        double th = jmax(f.rate, f.rate_k);
        contracted = f.stepnorm > 0.0 && th < 1.0 && th / (1.0 - th) * f.stepnorm <= q.kappa * xtol;
    }
    const bool step_conv = past_min && f.stepnorm > 0.0 && f.stepnorm <= xtol;
    const bool zero_step = s.code == kLsqSuccess && s.stepnorm == 0.0 && past_min;
    const bool roundoff = q.roundoff && past_min && f.rate >= 1.0 && f.rate_k < 1.0;
    if (contracted || (steptested && (step_conv || zero_step || roundoff))) {
        if (!refreshed) diag_pass<N>(f);
        if (stationary_dg<N>(f, q, f.it, f.fnorm, f.gr, f.jn, f.dg_exact, false, f.fnorm)) {
            f.code = kLsqSuccess;
            f.stationary = true;
            return true;
        }
        f.code = contracted ? kLsqSuccess : kLsqStalled;
        f.contracted = contracted;
        return true;
    }
    if (s.code == kLsqStalled) {
        f.stalls += 1;
        if (f.stalls >= q.max_stalls) {
            f.code = kLsqStalled;
            return true;
        }
    } else {
        f.stalls = 0;
    }
    if (f.it >= q.max_iterations) {
        f.code = kLsqMaxIters;
        return true;
    }
    return false;
}

struct LsqArgs {
    LsqParams q;
    double* X;               // nfit × 8
    const int32_t* m;
    const int64_t* y_offset;
    const double* Y;
    gs_status_f64* S;
    const int32_t* list;     // every fit, ordered by (K, m)
    const uint32_t* bstart;  // the buckets' starts in `list` (5 entries), written by lsq_scan
    const uint32_t* bad;     // the count of fits that fail validation, written by lsq_count
    int bucket;              // this bucket, K − 1
    uint32_t* queue;         // this bucket's queue head
    double* slots;           // this bucket's lane slots (Slot<N>::kSize entries per lane)
};

template <int N> GS_DEV void fit_load(Fit<N>& f, const LsqArgs& a, int64_t idx) {
    f.idx = idx;
    f.row.m = a.m[idx];
    f.row.y = a.Y + a.y_offset[idx];
    const double* xp = a.X + 8 * idx;
#pragma unroll
    for (int k = 0; k < N; ++k) f.x[k] = xp[k];
}

template <int N> GS_DEV void fit_store(Fit<N>& f, const LsqArgs& a) {
    // the reported gradratio with the reference's ‖J‖_F fold
    if (f.has_dg && !f.dg_exact) {
        double p[N];
        dg_point<N>(f, f.dg_at_xk, p);
        diag_exact<N>(f.row, p, f.dg_at_xk ? f.fnorm_k : f.fnorm, f.gr, f.jn);
        f.dg_exact = true;
    }
    double* xp = a.X + 8 * f.idx;
#pragma unroll
    for (int k = 0; k < N; ++k) xp[k] = f.x[k];
    uint64_t* w = reinterpret_cast<uint64_t*>(a.S + f.idx);
    w[0] = uint64_t(f.code) | (uint64_t(uint32_t(f.it)) << 32);
    w[1] = __builtin_bit_cast(uint64_t, f.fnorm);
    w[2] = __builtin_bit_cast(uint64_t, f.stepnorm);
    w[3] = uint64_t(uint32_t(f.failures)) | (uint64_t(f.stationary ? 1 : 0) << 40);
    w[4] = uint64_t(uint32_t(-1));
    w[5] = __builtin_bit_cast(uint64_t, dnan());
    w[6] = __builtin_bit_cast(uint64_t, f.gr);
    w[7] = __builtin_bit_cast(uint64_t, f.rate);
    w[8] = uint64_t(f.contracted ? 1 : 0);
}

// The persistent kernel of one bucket (fits with n = N). Every lane loops over LM outer steps; a lane
// without a fit takes the next one from the bucket's queue (one atomic per wave per refill), and the
// wave leaves the loop once the queue is empty and every lane is idle.
// The n = 6 kernel is held to 256 registers (2 waves per SIMD, about 60 words of its state spilled
// to scratch): measured 130 against 135-138 ms per step of the 262144-fit batch at 1 wave per SIMD
// with 314 registers. The n = 8 kernel stays at 1 wave per SIMD: held to 256 registers it spills
// about 230 words and runs no faster.
template <int N, bool SCALED>
__global__ __launch_bounds__(256, N == 6 ? 2 : 1) void lsq_kernel(const LsqArgs a) {
    // the bucket's range comes from the device-side sort, so the launch does not depend on the data;
    // a batch with an invalid fit solves nothing (the host reports the error after the launch)
    if (*a.bad != 0) return;
    const uint32_t lo = a.bstart[a.bucket];
    const uint64_t count = a.bstart[a.bucket + 1] - lo;
    const int32_t* list = a.list + lo;
    Fit<N> f;
    const int64_t lanes = int64_t(gridDim.x) * blockDim.x;
    f.slot.p = a.slots + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    f.slot.stride = lanes;
    bool active = false, exhausted = false;
    const uint64_t lane_lt = (uint64_t(1) << (threadIdx.x & 63)) - 1;
    while (true) {
        // refill the idle lanes (a fit may end at its start, so repeat while any lane is idle)
        while (true) {
            const bool need = !active && !exhausted;
            const uint64_t ballot = __ballot(need);
            if (ballot == 0) break;
            const uint32_t cnt = __popcll(ballot);
            uint32_t base = 0;
            const int leader = __ffsll(static_cast<long long>(ballot)) - 1;
            if ((threadIdx.x & 63) == leader) base = atomicAdd(a.queue, cnt);
            base = __shfl(base, leader);
            if (need) {
                const uint64_t slot = uint64_t(base) + __popcll(ballot & lane_lt);
                if (slot >= count) {
                    exhausted = true;
                } else {
                    fit_load<N>(f, a, list[slot]);
                    if (fit_start<N>(f, a.q))
                        fit_store<N>(f, a);
                    else
                        active = true;
                }
            }
        }
        if (!__any(active)) break;
        if (active && fit_iterate<N, SCALED>(f, a.q)) {
            fit_store<N>(f, a);
            active = false;
        }
    }
}

// ---- the bucketing: a counting sort of the fits over (K, m) bins ----------------------------------

__global__ void lsq_clear(uint32_t* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}

__global__ void lsq_count(const int32_t* K, const int32_t* m, const int64_t* y_offset, int64_t ny, int64_t nfit,
                          uint32_t* bins, uint32_t* bad) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nfit; i += int64_t(gridDim.x) * blockDim.x) {
        const int k = K[i], mm = m[i];
        const int64_t off = y_offset[i];
        if (k < 1 || k > 4 || mm < 2 * k || mm > kMaxM || off < 0 || off > ny - mm) {
            atomicAdd(bad, 1u);
            continue;
        }
        atomicAdd(&bins[(k - 1) * (kMaxM + 1) + mm], 1u);
    }
}

// Exclusive prefix sum of the bins (one workgroup; the bins are few) and the per-bucket starts.
__global__ __launch_bounds__(1024) void lsq_scan(const uint32_t* bins, uint32_t* cursor, uint32_t* bucket_start) {
    __shared__ uint32_t part[1024];
    const int per = (kBins + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = lo + per < kBins ? lo + per : kBins;
    uint32_t s = 0;
    for (int b = lo; b < hi; ++b) s += bins[b];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < 1024; ++t) {
            uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int b = lo; b < hi; ++b) {
        cursor[b] = run;
        if (b % (kMaxM + 1) == 0) bucket_start[b / (kMaxM + 1)] = run;
        run += bins[b];
    }
    if (threadIdx.x == 1023) bucket_start[4] = run;
}

__global__ void lsq_scatter(const int32_t* K, const int32_t* m, int64_t nfit, uint32_t* cursor, int32_t* list) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nfit; i += int64_t(gridDim.x) * blockDim.x) {
        const int k = K[i], mm = m[i];
        if (k < 1 || k > 4 || mm < 2 * k || mm > kMaxM) continue;
        const uint32_t pos = atomicAdd(&cursor[(k - 1) * (kMaxM + 1) + mm], 1u);
        list[pos] = int32_t(i);
    }
}

// ---- workspace and launch shapes -------------------------------------------------------------------

struct LsqWorkspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_lsq_mutex;
LsqWorkspace g_lsq_ws[64][2];  // per device: [0] bins, cursors and the ordered list; [1] lane slots

// The launch shape of each bucket: threads per workgroup and resident workgroups per CU (0: as many
// as the kernel's occupancy allows). The kernels are persistent, so the grid is the shape's share of
// the machine (capped by the batch), not the bucket's size.
struct LsqShape {
    int threads = 64;
    int blocks_per_cu = 0;
};
LsqShape g_lsq_shape[4] = {{64, 0}, {64, 0}, {64, 0}, {64, 0}};
bool g_lsq_graph_on = true;

// Library streams (default priority) on which the buckets run concurrently, one set per device, and
// the stream the launch sequence is captured on.
struct LsqStreams {
    hipStream_t s[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t cap = nullptr;
    hipEvent_t start = nullptr;
    hipEvent_t done[4] = {nullptr, nullptr, nullptr, nullptr};
};
LsqStreams g_lsq_streams[64];

int lsq_streams(LsqStreams** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_lsq_mutex);
    LsqStreams& L = g_lsq_streams[dev];
    if (!L.start) {
        for (int b = 0; b < 4; ++b) {
            if (hipStreamCreateWithFlags(&L.s[b], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&L.done[b], hipEventDisableTiming) != hipSuccess)
                return set_error(GS_EHIP, "stream creation failed");
        }
        if (hipStreamCreateWithFlags(&L.cap, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.start, hipEventDisableTiming) != hipSuccess)
            return set_error(GS_EHIP, "event creation failed");
    }
    *out = &L;
    return GS_OK;
}

int lsq_workspace(int which, size_t bytes, void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_lsq_mutex);
    LsqWorkspace& w = g_lsq_ws[dev][which];
    if (w.bytes < bytes) {
        if (w.ptr) (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
        if (hipMalloc(&w.ptr, bytes) != hipSuccess) return set_error(GS_EHIP, "hipMalloc of the least-squares workspace failed");
        w.bytes = bytes;
    }
    *out = w.ptr;
    return GS_OK;
}

template <int N, bool SCALED> const void* bucket_kernel() {
    return reinterpret_cast<const void*>(&lsq_kernel<N, SCALED>);
}
template <bool SCALED> const void* bucket_kernel_of(int b) {
    switch (b) {
        case 0: return bucket_kernel<2, SCALED>();
        case 1: return bucket_kernel<4, SCALED>();
        case 2: return bucket_kernel<6, SCALED>();
        default: return bucket_kernel<8, SCALED>();
    }
}
const void* bucket_kernel_of(int b, bool scaled) { return scaled ? bucket_kernel_of<true>(b) : bucket_kernel_of<false>(b); }

// Workgroups of bucket b under its shape for a batch of nfit fits: the resident workgroups the shape
// asks for (the occupancy's when blocks_per_cu is 0), and no more lanes than fits.
int lsq_bucket_blocks(int b, int64_t nfit, bool scaled, int64_t* out) {
    const LsqShape& sh = g_lsq_shape[b];
    int dev = 0, cus = 256, per_cu = sh.blocks_per_cu;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(GS_EHIP, "hipGetDevice failed");
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (per_cu == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bucket_kernel_of(b, scaled), sh.threads, 0) !=
                            hipSuccess || per_cu < 1))
        per_cu = 1;
    int64_t blocks = int64_t(per_cu) * cus;
    const int64_t cap = (nfit + sh.threads - 1) / sh.threads;
    *out = blocks < cap ? blocks : cap;
    return GS_OK;
}

// The captured launch sequence of one device, and what it was captured for.
struct LsqGraph {
    hipGraphExec_t exec = nullptr;
    std::string key;
};
LsqGraph g_lsq_graph[64];

// The whole launch sequence on `st`: clear, count, scan, scatter, then the four bucket kernels on the
// library's streams, forked from and joined back to `st` with events.
int lsq_enqueue(hipStream_t st, const LsqStreams& L, const int32_t* K, const int32_t* m, const int64_t* y_offset,
                int64_t ny, int64_t nfit, uint32_t* bins, int64_t words, LsqArgs a, const int64_t (&blocks)[4],
                double* const (&slots)[4]) {
    const size_t nb = size_t(kBins);
    uint32_t* cursor = bins + nb;
    uint32_t* bstart = cursor + nb;
    uint32_t* queues = bstart + 5;
    uint32_t* bad = queues + 4;
    int32_t* list = reinterpret_cast<int32_t*>(bad + 1);
    hipLaunchKernelGGL(lsq_clear, dim3(64), dim3(256), 0, st, bins, words);
    const int gb = int((nfit + 255) / 256 < 4096 ? (nfit + 255) / 256 : 4096);
    hipLaunchKernelGGL(lsq_count, dim3(gb), dim3(256), 0, st, K, m, y_offset, ny, nfit, bins, bad);
    hipLaunchKernelGGL(lsq_scan, dim3(1), dim3(1024), 0, st, bins, cursor, bstart);
    hipLaunchKernelGGL(lsq_scatter, dim3(gb), dim3(256), 0, st, K, m, nfit, cursor, list);
    a.list = list;
    a.bstart = bstart;
    a.bad = bad;
    if (hipEventRecord(L.start, st) != hipSuccess) return set_error(GS_EHIP, "hipEventRecord failed");
    for (int b = 0; b < 4; ++b) {
        a.bucket = b;
        a.queue = queues + b;
        a.slots = slots[b];
        hipStream_t sb = L.s[b];
        if (hipStreamWaitEvent(sb, L.start, 0) != hipSuccess) return set_error(GS_EHIP, "hipStreamWaitEvent failed");
        const dim3 grid{unsigned(blocks[b]), 1, 1}, block{unsigned(g_lsq_shape[b].threads), 1, 1};
        if (a.q.scaled) {
            switch (b) {
                case 0: hipLaunchKernelGGL((lsq_kernel<2, true>), grid, block, 0, sb, a); break;
                case 1: hipLaunchKernelGGL((lsq_kernel<4, true>), grid, block, 0, sb, a); break;
                case 2: hipLaunchKernelGGL((lsq_kernel<6, true>), grid, block, 0, sb, a); break;
                case 3: hipLaunchKernelGGL((lsq_kernel<8, true>), grid, block, 0, sb, a); break;
            }
        } else {
            switch (b) {
                case 0: hipLaunchKernelGGL((lsq_kernel<2, false>), grid, block, 0, sb, a); break;
                case 1: hipLaunchKernelGGL((lsq_kernel<4, false>), grid, block, 0, sb, a); break;
                case 2: hipLaunchKernelGGL((lsq_kernel<6, false>), grid, block, 0, sb, a); break;
                case 3: hipLaunchKernelGGL((lsq_kernel<8, false>), grid, block, 0, sb, a); break;
            }
        }
        if (hipEventRecord(L.done[b], sb) != hipSuccess || hipStreamWaitEvent(st, L.done[b], 0) != hipSuccess)
            return set_error(GS_EHIP, "joining the bucket streams failed");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(GS_EHIP, std::string("least-squares launch failed: ") + hipGetErrorString(e));
    return GS_OK;
}

}  // namespace

// Free the captured graphs, the library streams and events, and the workspaces of every device
// (gs_shutdown; no call may be in flight). The next call creates them again.
void release_lsq() {
    std::lock_guard<std::mutex> lock(g_lsq_mutex);
    for (int d = 0; d < 64; ++d) {
        LsqGraph& G = g_lsq_graph[d];
        LsqStreams& L = g_lsq_streams[d];
        const bool used = G.exec || L.start || g_lsq_ws[d][0].ptr || g_lsq_ws[d][1].ptr;
        if (!used || hipSetDevice(d) != hipSuccess) continue;
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        G.exec = nullptr;
        G.key.clear();
        for (int b = 0; b < 4; ++b) {
            if (L.s[b]) (void)hipStreamDestroy(L.s[b]);
            if (L.done[b]) (void)hipEventDestroy(L.done[b]);
        }
        if (L.cap) (void)hipStreamDestroy(L.cap);
        if (L.start) (void)hipEventDestroy(L.start);
        L = LsqStreams{};
        for (auto& w : g_lsq_ws[d]) {
            if (w.ptr) (void)hipFree(w.ptr);
            w = LsqWorkspace{};
        }
    }
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_solve_expsum_lsq_f64(const gs_method* method, int64_t nfit, double* X, const int32_t* K, const int32_t* m,
                            const int64_t* y_offset, const double* Y, int64_t ny, gs_status_f64* S,
                            const gs_options_f64* options, void* stream) {
    if (!method || !options) return set_error(GS_EINVAL, "method and options must not be NULL");
    if (method->globalization != GS_LEVENBERG_MARQUARDT)
        return set_error(GS_EFAMILY, "the least-squares batch takes Newton(; globalization = LevenbergMarquardt())");
    if (method->reuse != 0) return set_error(GS_EINVAL, "Newton with a LevenbergMarquardt globalization takes no reuse");
    // the LevenbergMarquardt constructor's checks (levenbergmarquardt.jl:60-95); they also bound the
    // damping loop of a step, which ends once μ passes μ_max
    {
        const double mu0 = method->lm_mu0, lo = method->lm_mu_min, hi = method->lm_mu_max;
        const bool fin = std::isfinite(mu0) && std::isfinite(lo) && std::isfinite(hi);
        if (!(fin && 0.0 < lo && lo <= mu0 && mu0 <= hi))
            return set_error(GS_EINVAL, "LevenbergMarquardt requires finite damping bounds 0 < μ_min ≤ μ₀ ≤ μ_max < Inf");
        if (!(std::isfinite(method->lm_mu_decrease) && 0.0 < method->lm_mu_decrease && method->lm_mu_decrease < 1.0))
            return set_error(GS_EINVAL, "LevenbergMarquardt requires 0 < μ_decrease < 1");
        if (!(std::isfinite(method->lm_mu_increase) && method->lm_mu_increase > 1.0))
            return set_error(GS_EINVAL, "LevenbergMarquardt requires μ_increase > 1");
        if (!(std::isfinite(method->lm_rho_accept) && 0.0 <= method->lm_rho_accept && method->lm_rho_accept < 1.0))
            return set_error(GS_EINVAL, "LevenbergMarquardt requires 0 ≤ ρ_accept < 1");
        if (method->lm_scaling != GS_LM_SCALING_NONE && method->lm_scaling != GS_LM_SCALING_DIAGONAL)
            return set_error(GS_EINVAL, "LevenbergMarquardt requires scaling ∈ (:none, :diagonal)");
        if (options->max_iterations < 0 || options->min_iterations < 0 || options->max_stalls < 1)
            return set_error(GS_EINVAL, "max_iterations and min_iterations must be ≥ 0 and max_stalls ≥ 1");
    }
    if (nfit < 0 || nfit > 0x7fffffffLL) return set_error(GS_EINVAL, "nfit must be in 0:2^31-1");
    if (nfit == 0) return GS_OK;
    if (ny < 0) return set_error(GS_EINVAL, "ny must be ≥ 0");
    if (!X || !K || !m || !y_offset || !Y || !S) return set_error(GS_EINVAL, "X, K, m, y_offset, Y and S must not be NULL");
    if ((reinterpret_cast<uintptr_t>(X) % 8) || (reinterpret_cast<uintptr_t>(S) % 8))
        return set_error(GS_EINVAL, "X and S must be 8-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    // workspace: bins, cursors, 5 bucket starts, 4 queues, 1 error count, the ordered list
    const size_t nb = size_t(kBins);
    const int64_t words = int64_t(2 * nb + 5 + 4 + 1);
    void* ws = nullptr;
    int rc = lsq_workspace(0, size_t(words) * 4 + size_t(nfit) * 4, &ws);
    if (rc) return rc;
    // the lane slots of the four buckets (Slot<N>::kSize doubles per lane)
    int64_t blocks[4];
    size_t slot_off[4], slot_total = 0;
    for (int b = 0; b < 4; ++b) {
        const int n = 2 * (b + 1);
        if ((rc = lsq_bucket_blocks(b, nfit, method->lm_scaling == GS_LM_SCALING_DIAGONAL, &blocks[b]))) return rc;
        slot_off[b] = slot_total;
        slot_total += size_t(blocks[b] * g_lsq_shape[b].threads) * size_t(n * (n + 1) / 2 + 2 * n);
    }
    void* slot_ws = nullptr;
    if ((rc = lsq_workspace(1, slot_total * 8, &slot_ws))) return rc;
    double* slots[4];
    for (int b = 0; b < 4; ++b) slots[b] = static_cast<double*>(slot_ws) + slot_off[b];
    LsqArgs a;
    std::memset(&a, 0, sizeof(a));
    LsqParams& q = a.q;
    q.f_abstol = options->f_abstol;
    q.f_reltol = options->f_reltol;
    q.x_abstol = options->x_abstol;
    q.x_reltol = options->x_reltol;
    q.kappa = options->kappa;
    q.g_tol = options->g_tol;
    q.min_iterations = options->min_iterations;
    q.max_iterations = options->max_iterations;
    q.max_stalls = options->max_stalls;
    q.norm_max = options->norm == GS_NORM_MAX;
    q.roundoff = options->roundoff != 0;
    q.rate_test = options->rate_test != 0;
    q.mu0 = method->lm_mu0;
    q.mu_min_eff = method->lm_mu_min > 4.0 * 2.220446049250313e-16 ? method->lm_mu_min : 4.0 * 2.220446049250313e-16;
    q.mu_max = method->lm_mu_max;
    q.mu_dec = method->lm_mu_decrease;
    q.mu_inc = method->lm_mu_increase;
    q.rho_accept = method->lm_rho_accept;
    q.scaled = method->lm_scaling == GS_LM_SCALING_DIAGONAL;
    a.X = X;
    a.m = m;
    a.y_offset = y_offset;
    a.Y = Y;
    a.S = S;
    LsqStreams* L = nullptr;
    if ((rc = lsq_streams(&L))) return rc;
    uint32_t* bins = static_cast<uint32_t*>(ws);
    const uint32_t* bad = bins + 2 * nb + 5 + 4;
    if (g_lsq_graph_on) {
        // one captured graph per device, replayed while the call's pointers, sizes, parameters and
        // shapes are those it was captured with
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
        const void* ptrs[8] = {X, K, m, y_offset, Y, S, ws, slot_ws};
        std::string key;
        key.append(reinterpret_cast<const char*>(&nfit), sizeof(nfit));
        key.append(reinterpret_cast<const char*>(&ny), sizeof(ny));
        key.append(reinterpret_cast<const char*>(ptrs), sizeof(ptrs));
        key.append(reinterpret_cast<const char*>(&q), sizeof(q));
        key.append(reinterpret_cast<const char*>(blocks), sizeof(blocks));
        key.append(reinterpret_cast<const char*>(g_lsq_shape), sizeof(g_lsq_shape));
        std::lock_guard<std::mutex> lock(g_lsq_mutex);
        LsqGraph& G = g_lsq_graph[dev];
        if (!G.exec || G.key != key) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            G.exec = nullptr;
            G.key.clear();
            if (hipStreamBeginCapture(L->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return set_error(GS_EHIP, "hipStreamBeginCapture failed");
            rc = lsq_enqueue(L->cap, *L, K, m, y_offset, ny, nfit, bins, words, a, blocks, slots);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(L->cap, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ec != hipSuccess || !graph) return set_error(GS_EHIP, "hipStreamEndCapture failed");
            const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                G.exec = nullptr;
                return set_error(GS_EHIP, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei));
            }
            G.key = key;
        }
        if (hipGraphLaunch(G.exec, st) != hipSuccess) return set_error(GS_EHIP, "hipGraphLaunch failed");
    } else {
        if ((rc = lsq_enqueue(st, *L, K, m, y_offset, ny, nfit, bins, words, a, blocks, slots))) return rc;
    }
    uint32_t host_bad = 0;
    if (hipMemcpyAsync(&host_bad, bad, 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return set_error(GS_EHIP, "least-squares status copy failed");
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(GS_EHIP, std::string("hipStreamSynchronize failed: ") + hipGetErrorString(e));
    if (host_bad != 0)
        return set_error(GS_EINVAL, std::to_string(host_bad) + " fits have K outside 1:4, m outside 2K:" +
                                        std::to_string(kMaxM) + " or data outside Y");
    return GS_OK;
}

// The launch shape of bucket k (n = 2k): threads per workgroup (64, 128, 192 or 256) and resident
// workgroups per CU (0: the kernel's occupancy, else 1..64), as the package's autotuner sets it.
int gs_lsq_set_shape(int k, int threads, int blocks_per_cu) {
    if (k < 1 || k > 4) return set_error(GS_EINVAL, "k must be in 1:4");
    if (threads < 64 || threads > 256 || threads % 64) return set_error(GS_EINVAL, "threads must be 64, 128, 192 or 256");
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return set_error(GS_EINVAL, "blocks_per_cu must be in 0:64");
    std::lock_guard<std::mutex> lock(g_lsq_mutex);
    g_lsq_shape[k - 1].threads = threads;
    g_lsq_shape[k - 1].blocks_per_cu = blocks_per_cu;
    return GS_OK;
}

// The shape of bucket k and, for a batch of nfit fits, its workgroups (out[0..2]; the occupancy of
// the scaling = :none kernels where the shape leaves it to the occupancy).
int gs_lsq_get_shape(int k, int64_t nfit, int64_t out[3]) {
    if (k < 1 || k > 4 || !out) return set_error(GS_EINVAL, "k must be in 1:4 and out not NULL");
    int64_t blocks = 0;
    const int rc = lsq_bucket_blocks(k - 1, nfit < 1 ? 1 : nfit, false, &blocks);
    if (rc) return rc;
    out[0] = g_lsq_shape[k - 1].threads;
    out[1] = g_lsq_shape[k - 1].blocks_per_cu;
    out[2] = blocks;
    return GS_OK;
}

// Capture the least-squares launch sequence into a hipGraph and replay it (1, the default) or
// launch it eagerly (0).
int gs_lsq_set_graph(int enable) {
    std::lock_guard<std::mutex> lock(g_lsq_mutex);
    g_lsq_graph_on = enable != 0;
    return GS_OK;
}

}  // extern "C"
