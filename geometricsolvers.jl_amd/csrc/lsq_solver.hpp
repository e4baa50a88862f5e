// The least-squares batch: Newton + LevenbergMarquardt on exponential-sum fits, one lane per fit, in
// a persistent kernel whose lanes refill from a work queue.
//
// Per fit this is the reference's R1 `solve!!` on a LeastSquaresProblem (src/nonlinear/solver.jl:
// 148-189, 277-378) with the LevenbergMarquardt step (src/nonlinear/newton.jl:1876-1971) on the
// stacked Givens reduction of [J | −F] (src/globalization/levenbergmarquardt.jl:359-482, 575-584).
// The fit's m rows are never stored: every pass over them recomputes the row (t_i, its exponentials,
// F_i and J_i) and streams it — into the Givens reduction in the reference's row order, into the
// left-fold sums of the gain ratio, or into the diagnostics ‖Jᴴ F‖ and ‖J‖_F. The triangular
// factor, its right-hand side, the iterate and the solve's status live in VGPRs.
//
// A lane whose fit has ended takes the next fit of its bucket from the queue at the top of the
// next outer step, so a wave is never held by the slowest of its 64 fits for longer than one step.
#pragma once

#include "device_common.hpp"

namespace gs {

// ReturnCode values (src/base/returncodes.jl:14) that the least-squares solve can end with.
enum : uint8_t { kLsqSuccess = 0, kLsqStalled = 1, kLsqMaxIters = 2, kLsqNonfinite = 4 };

struct LsqParams {
    // Options
    double f_abstol, f_reltol, x_abstol, x_reltol, kappa, g_tol;
    int32_t min_iterations, max_iterations, max_stalls, norm_max, roundoff, rate_test;
    // LevenbergMarquardt
    double mu0, mu_min_eff, mu_max, mu_dec, mu_inc, rho_accept;
    int32_t scaled;  // scaling = :diagonal (the kernels of lsq_kernel<N, true>)
    int32_t pad_;
};

// LinearAlgebra's givensAlgorithm (the LAPACK DLARTG scheme): r = √(f² + g²), c = f/r, s = g/r,
// power-of-two rescaling outside [2^-537, 2^537], c ≥ 0 where |f| > |g|.
//
// Written without branches, with the same results bit for bit: for finite f, g the reference's
// rescaling loops run once (scale ≥ 2^537: one factor 2^-537 brings scale below 2^487; scale ≤
// 2^-537: one factor 2^537 lifts it above 2^-537 unless scale is 2^-1074 itself, which takes two);
// every factor is exact, and so is every undo (r · 2^537, or r · 2^-537 once or twice). An infinite
// scale makes the reference loop 20 times, with the same r = ±Inf, c = NaN and s = ±0 (or NaN) as
// one factor gives; a NaN scale takes the direct path in both.
GS_DEV void givens(double f, double g, double& c, double& s, double& r) {
    constexpr double safmn2 = 0x1p-537, safmx2 = 0x1p537;
    const double af = __builtin_fabs(f), ag = __builtin_fabs(g);
    const double scale = jmax(af, ag);
    const bool big = scale >= safmx2, small = scale <= safmn2;
    const double m1 = big ? safmn2 : (small ? safmx2 : 1.0);
    const double m2 = (small && scale == 0x1p-1074) ? safmx2 : 1.0;
    const double f1 = (f * m1) * m2, g1 = (g * m1) * m2;
    double rr = __builtin_sqrt(f1 * f1 + g1 * g1);
    double cc = f1 / rr, ss = g1 / rr;
    const double u1 = big ? safmx2 : (small ? safmn2 : 1.0);
    const double u2 = (small && scale == 0x1p-1074) ? safmn2 : 1.0;
    rr = (rr * u1) * u2;
    const bool flip = af > ag && cc < 0.0;
    cc = flip ? -cc : cc;
    ss = flip ? -ss : ss;
    rr = flip ? -rr : rr;
    // g == 0 → (1, 0, f); f == 0 → (0, 1, g)
    c = g == 0.0 ? 1.0 : (f == 0.0 ? 0.0 : cc);
    s = g == 0.0 ? 0.0 : (f == 0.0 ? 1.0 : ss);
    r = g == 0.0 ? f : (f == 0.0 ? g : rr);
}

// givens() where it needs none of its special handling: f ≠ 0, g ≠ 0 and max(|f|, |g|) in
// [2^-383, 2^511]. There the rescaling factors are 1, neither zero case applies and f² + g² lies in
// [2^-766, 2^1023], where sqrt_inrange has the bits of the square root, so the results are givens()'s
// bit for bit, without its selects.
GS_DEV bool givens_plain_ok(double f, double g) {
    const double scale = jmax(__builtin_fabs(f), __builtin_fabs(g));  // NaN fails the range test
    return f != 0.0 && g != 0.0 && scale >= 0x1p-383 && scale <= 0x1p511;
}
GS_DEV void givens_plain(double f, double g, double& c, double& s, double& r) {
    const double rr = sqrt_inrange(f * f + g * g);
    const double cc = f / rr, ss = g / rr;
    const bool flip = __builtin_fabs(f) > __builtin_fabs(g) && cc < 0.0;
    c = flip ? -cc : cc;
    s = flip ? -ss : ss;
    r = flip ? -rr : rr;
}

// The Taylor coefficients 1/j!, j = 13 … 2, rounded to double. Read from constant memory inside the
// row loops, they stay in scalar registers (the FMA takes one scalar operand) instead of occupying
// VGPRs for the life of the kernel.
__constant__ double kExpTaylor[12] = {
    0x1.6124613a86d09p-33, 0x1.1eed8eff8d898p-29, 0x1.ae64567f544e4p-26, 0x1.27e4fb7789f5cp-22,
    0x1.71de3a556c734p-19, 0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-13, 0x1.6c16c16c16c17p-10,
    0x1.1111111111111p-7,  0x1.5555555555555p-5,  0x1.5555555555555p-3,  0.5};

// The exponential of the exponential-sum family, a fixed algorithm so that the device and the CPU
// oracle evaluate F and J to the same bits: e^x = 2^k e^r with k = rint(x / ln 2) and
// r = (x − k L1) − k L2 (Cody–Waite: L1 = ln 2 rounded to a multiple of 2^-41, so k L1 is exact for
// |k| < 2^12, and L2 = ln 2 − L1), e^r from its Taylor polynomial of degree 13 in Horner form with
// fused multiply-adds (|r| ≤ 0.3466: truncation below 2^-57 relative), then scaled by 2^k. Overflow
// above ln(2^1024), 0 below ln(2^-1075), NaN unchanged. Within 1 ulp of e^x.
// Branch-free: the polynomial is evaluated for every x and the special cases selected afterwards (a
// NaN x propagates through it; the overflow and underflow thresholds are tested on x).
GS_DEV double expsum_exp(double x) {
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-kd, 0x1.62e42fefa4000p-1, x);
    r = __builtin_fma(-kd, -0x1.8432a1b0e2634p-43, r);
    double p = kExpTaylor[0];
#pragma unroll
    for (int j = 1; j < 12; ++j) p = __builtin_fma(p, r, kExpTaylor[j]);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    // |kd| ≤ 1075 wherever the result is used; clamp so the conversion is defined everywhere
    const double kc = __builtin_fmin(__builtin_fmax(kd, -2000.0), 2000.0);
    const double e = __builtin_ldexp(p, int(kc));
    return x > 0x1.62e42fefa39efp+9 ? __builtin_inf() : (x < -0x1.74910d52d3052p+9 ? 0.0 : e);
}

// One fit of the exponential-sum family, x = (a_1..a_K, λ_1..λ_K), t_i = 5i/(m−1) (i from 0):
// F_i = (Σ_k a_k exp(−λ_k t_i)) − y_i, J_ik = exp(−λ_k t_i), J_i,K+k = ((−a_k) t_i) exp(−λ_k t_i).
template <int N> struct ExpSumRow {
    static constexpr int K = N / 2;
    const double* y;
    int m;
    GS_DEV double t(int i) const { return (5.0 * double(i)) / double(m - 1); }
    GS_DEV double F(const double (&x)[N], int i) const { return F(x, i, y[i]); }
    GS_DEV double FJ(const double (&x)[N], int i, double (&Ji)[N]) const { return FJ(x, i, y[i], Ji); }
    // the same with y_i given (read ahead by the caller's row loop)
    GS_DEV double F(const double (&x)[N], int i, double yi) const {
        const double ti = t(i);
        double s = x[0] * expsum_exp(-x[K] * ti);
#pragma unroll
        for (int k = 1; k < K; ++k) s = s + x[k] * expsum_exp(-x[K + k] * ti);
        return s - yi;
    }
    GS_DEV double FJ(const double (&x)[N], int i, double yi, double (&Ji)[N]) const {
        const double ti = t(i);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double e = expsum_exp(-x[K + k] * ti);
            s = k == 0 ? x[0] * e : s + x[k] * e;
            Ji[k] = e;
            Ji[K + k] = ((-x[k]) * ti) * e;
        }
        return s - yi;
    }
    // y_i of a row loop from row i on: the load of the next row's value is issued one row ahead, so
    // that its latency hides under the row's arithmetic (the index stays in 0 … m − 1)
    struct Ahead {
        const double* y;
        int last;
        double next;
        GS_DEV Ahead(const ExpSumRow& r, int i0) : y(r.y), last(r.m - 1), next(r.y[i0 < r.m - 1 ? i0 : r.m - 1]) {}
        GS_DEV double take(int i) {
            const double v = next;
            next = y[i + 1 < last ? i + 1 : last];
            return v;
        }
    };
    // J_ik alone (column k of row i)
    GS_DEV double J(const double (&x)[N], int i, int k) const {
        const double ti = t(i);
        if (k < K) return expsum_exp(-x[K + k] * ti);
        return ((-x[k - K]) * ti) * expsum_exp(-x[k] * ti);
    }
};

// Compile-time loop: fn(IC<i>{}) for i = B, …, E−1, as straight-line code (no loop for the unroller
// to give up on, so every array index stays a constant and the arrays stay in VGPRs).
template <int V> struct IC {
    static constexpr int value = V;
    constexpr operator int() const { return V; }
};
template <int B, int E, class Fn> GS_DEV void static_for(Fn&& fn) {
    if constexpr (B < E) {
        fn(IC<B>{});
        static_for<B + 1, E>(fn);
    }
}

// Rotate row `w` (with right-hand side wr) into the upper-triangular factor (R, c), starting at
// column K0 — the inner step of lm_qr_factor! (:402-419) and of lm_qr_solve! (:449-466).
template <int K0, int N> GS_DEV void rotate_in(double (&R)[N][N], double (&c)[N], double (&w)[N], double& wr) {
    static_for<K0, N>([&](auto kk) {
        constexpr int k = kk;
        const double b = w[k];
        if (__all(givens_plain_ok(R[k][k], b))) {
            // the common case of every lane of the wave: no zero entry, no rescaling
            double cs, sn, r;
            givens_plain(R[k][k], b, cs, sn, r);
            R[k][k] = r;
            w[k] = 0.0;
            static_for<k + 1, N>([&](auto ll) {
                constexpr int l = ll;
                const double u = R[k][l], v = w[l];
                R[k][l] = cs * u + sn * v;
                w[l] = -sn * u + cs * v;
            });
            const double ur = c[k];
            c[k] = cs * ur + sn * wr;
            wr = -sn * ur + cs * wr;
            return;
        }
        // a zero entry is skipped (the rotation is computed and not applied), without a branch
        const bool skip = b == 0.0;
        double cs, sn, r;
        givens(R[k][k], b, cs, sn, r);
        R[k][k] = skip ? R[k][k] : r;
        w[k] = 0.0;
        static_for<k + 1, N>([&](auto ll) {
            constexpr int l = ll;
            const double u = R[k][l], v = w[l];
            R[k][l] = skip ? u : cs * u + sn * v;
            w[l] = skip ? v : -sn * u + cs * v;
        });
        const double ur = c[k];
        c[k] = skip ? ur : cs * ur + sn * wr;
        wr = skip ? wr : -sn * ur + cs * wr;
    });
}

}  // namespace gs
