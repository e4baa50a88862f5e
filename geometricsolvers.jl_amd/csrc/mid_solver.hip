// libgsolve_hip, mid-size systems: one wave per system, the factor in LDS.
//
// The FPU-β Gauss–Legendre s = 3 stage system (n = 48, GS_FAM_FPU_GL3) under Newton with the Static
// globalization, in the working precision (static LU in FP64) or mixed (LUFactorization(Float32)
// with Refinement and promotion to FP64): the reference's R1 loop solve!! (src/nonlinear/solver.jl:
// 277-378) with newton_step!! (src/nonlinear/newton.jl:1604-1720) per system, as the lane-per-system
// kernels run it for n ≤ 8.
//
// Lane i < 48 owns entry i of every vector (in LDS) and row i of the factor; the scalar state of the
// solve is uniform across the wave. Sums that the reference folds left to right (norms, the
// matrix-vector products) are folded in that order: the norms by every lane over the LDS vector
// (broadcast reads), the products of the tridiagonal-block Jacobian by the lane of their row or
// column over its nonzero entries only (a structural zero adds an exact zero to the fold). The
// Jacobian itself is never stored: its entries are recomputed from the bond stiffnesses K.
//
// The FP64 factorisation is the reference's unblocked LU (first-largest pivot by a wave reduction,
// a/p multipliers, rank-1 updates without FMA), so the working-precision solve matches the oracle
// bit for bit. The FP32 factor of the mixed method is blocked (16-column panels): the panels and
// U12 are factored as the reference does, and the trailing update A22 −= L21 U12 runs on
// v_mfma_f32_16x16x4f32, whose sums round differently; the FP64 refinement corrects the step.
#include <hip/hip_runtime.h>

#include <string>

#include "mid_solver.hpp"

namespace gs {

int set_error(int code, const std::string& msg);

namespace {

constexpr int kN = 48;
constexpr int kSites = 16;

// Gauss–Legendre s = 3: c and A² (A the collocation matrix), rounded once from 50-digit values.
__constant__ double kGlC[3] = {0x1.cda042f0236e1p-4, 0x1.0000000000000p-1, 0x1.c64bf7a1fb924p-1};
__constant__ double kGlA2[9] = {0x1.6c16c16c16c17p-7, -0x1.0f9922995a9d1p-7, 0x1.ce741a1dacb5ep-9,
                                0x1.a370da4233f7dp-4, 0x1.c71c71c71c71cp-6,  -0x1.537f6b3fb1445p-8,
                                0x1.bfe2a15ea5befp-3, 0x1.4f8d7b6829931p-3, 0x1.6c16c16c16c17p-7};

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct alignas(16) MidShared {
    double A[kN * kN];  // the factor: FP64, or FP32 in its first half
    double x[kN], Fx[kN], d[kN], r[kN], t1[kN], t2[kN], t3[kN], bx[kN];
    float t32[kN];
    double qp[32];      // q, p of the system
    double Qv[kN];      // stage positions
    double g[kN];       // stage forces
    double K[3 * 17];   // bond stiffnesses at the Jacobian's point
    int piv[kN];
};

GS_DEV void sync() { __syncthreads(); }

GS_DEV int lane_id() { return int(threadIdx.x); }

GS_DEV double wave_fmax(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = __builtin_fmax(v, __shfl_xor(v, o));
    return v;
}

// Left folds over an LDS vector of 48, evaluated by every lane (uniform results).
GS_DEV double fold_norm2(const double* v) {
    double s = v[0] * v[0];
    for (int i = 1; i < kN; ++i) s = s + v[i] * v[i];
    return s;
}

GS_DEV double fold_rmax(const double* v) {
    double m = __builtin_fabs(v[0]);
    for (int i = 1; i < kN; ++i) m = jmax(m, __builtin_fabs(v[i]));
    return m;
}

// rnorm (reductions.jl:41-57).
GS_DEV double fold_rnorm(const double* v) {
    const double s = fold_norm2(v);
    if (__builtin_expect(isfin(s) && s >= double(kN) * kDblMin, 1)) return __builtin_sqrt(s);
    const double m = fold_rmax(v);
    if (m == 0.0 || !isfin(m)) return m;
    const int e = exponent_of(m);
    double t = ldexp2(v[0], -e);
    double acc = t * t;
    for (int i = 1; i < kN; ++i) {
        t = ldexp2(v[i], -e);
        acc = acc + t * t;
    }
    return ldexp2(__builtin_sqrt(acc), e);
}

// The exponent of the largest |entry| of a vector (matrix_exponent, factors.jl:826-843), over the
// wave: 0 for a zero vector or one with a non-finite entry.
GS_DEV int vec_exponent(const double* v) {
    const int l = lane_id();
    const double a = l < kN ? __builtin_fabs(v[l]) : 0.0;
    const bool nan_any = __any(l < kN && isnan_(a));
    const double m = wave_fmax(a);
    if (nan_any || !isfin(m) || m == 0.0) return 0;
    const int e = exponent_of(m);
    return e < -kEmax64 ? -kEmax64 : (e > kEmax64 ? kEmax64 : e);
}

// rhs_exponent (factors.jl:880-885) for n = 48.
GS_DEV int vec_rhs_exponent(const double* r, int e) {
    const int m = vec_exponent(r);
    const int lo = e < 0 ? e : 0;
    const int c = m < lo ? lo : (m > 0 ? 0 : m);
    const int alt = m - (kEmax64 - headroom(kN));
    return c > alt ? c : alt;
}

// ---- the family -----------------------------------------------------------------------------------

struct Fpu {
    double h, beta, h2;
    // h c_a and h² A²_ab, formed where they are used (a run-time index into constant memory)
    GS_DEV double hc(int a) const { return h * kGlC[a]; }
    GS_DEV double coef(int a, int b) const { return h2 * kGlA2[a * 3 + b]; }
};

GS_DEV Fpu fpu_of(const double* shared) {
    Fpu f;
    f.h = shared[0];
    f.beta = shared[1];
    f.h2 = f.h * f.h;
    return f;
}

// Stage positions Q = q + z (into Qv).
GS_DEV void stages(MidShared& s, const double* z) {
    const int l = lane_id();
    if (l < kN) s.Qv[l] = s.qp[l % kSites] + z[l];
    sync();
}

// r_s = Q_s − Q_{s−1} of stage b (bond s = 0..16, the ends at 0).
GS_DEV double bond(const MidShared& s, int b, int k) {
    const double hi = k < kSites ? s.Qv[b * kSites + k] : 0.0;
    const double lo = k > 0 ? s.Qv[b * kSites + k - 1] : 0.0;
    return hi - lo;
}

// F(z) into out (z and out in LDS, out ≠ z).
GS_DEV void eval_F(MidShared& s, const Fpu& f, const double* z, double* out) {
    stages(s, z);
    const int l = lane_id();
    if (l < kN) {
        const int b = l / kSites, m = l % kSites;
        const double rl = bond(s, b, m), rr = bond(s, b, m + 1);
        const double dl = rl + f.beta * ((rl * rl) * rl), dr = rr + f.beta * ((rr * rr) * rr);
        s.g[l] = dr - dl;
    }
    sync();
    if (l < kN) {
        const int a = l / kSites, m = l % kSites;
        double t = kGlA2[a * 3 + 0] * s.g[m];
        t = t + kGlA2[a * 3 + 1] * s.g[kSites + m];
        t = t + kGlA2[a * 3 + 2] * s.g[2 * kSites + m];
        out[l] = (z[l] - f.hc(a) * s.qp[kSites + m]) - f.h2 * t;
    }
    sync();
}

// The bond stiffnesses K_b,s = 1 + 3β r² at z (the Jacobian's point).
GS_DEV void eval_K(MidShared& s, const Fpu& f, const double* z) {
    stages(s, z);
    const int l = lane_id();
    if (l < 3 * 17) {
        const int b = l / 17, k = l % 17;
        const double rv = bond(s, b, k);
        s.K[l] = 1.0 + (3.0 * f.beta) * (rv * rv);
    }
    sync();
}

// J_ij (unregularised): δ − (h² A²_ab) H^b_mk.
GS_DEV double jentry(const MidShared& s, const Fpu& f, int i, int j) {
    const int a = i / kSites, m = i % kSites, b = j / kSites, k = j % kSites;
    double H = 0.0;
    if (k == m) H = -(s.K[b * 17 + m] + s.K[b * 17 + m + 1]);
    else if (k + 1 == m) H = s.K[b * 17 + m];
    else if (k == m + 1) H = s.K[b * 17 + m + 1];
    return (i == j ? 1.0 : 0.0) - f.coef(a, b) * H;
}

// Jreg_ij = J_ij + μ δ_ij.
GS_DEV double jreg(const MidShared& s, const Fpu& f, double mu, int i, int j) {
    const double v = jentry(s, f, i, j);
    return (i == j && mu != 0.0) ? v + mu : v;
}

// y = Jreg v (row folds of StaticArrays' muladd chain over the nonzero columns).
GS_DEV void jmatvec(MidShared& s, const Fpu& f, double mu, const double* v, double* y) {
    const int l = lane_id();
    if (l < kN) {
        const int m = l % kSites;
        double acc = 0.0;
        bool first = true;
        for (int b = 0; b < 3; ++b) {
            for (int dk = -1; dk <= 1; ++dk) {
                const int k = m + dk;
                if (k < 0 || k >= kSites) continue;
                const int j = b * kSites + k;
                const double Jij = jreg(s, f, mu, l, j);
                acc = first ? Jij * v[j] : __builtin_fma(Jij, v[j], acc);
                first = false;
            }
        }
        y[l] = acc;
    }
    sync();
}

// ‖J‖_F: the reference's left fold over the entries in column-major order (zeros add nothing).
GS_DEV double jnorm_fold(const MidShared& s, const Fpu& f) {
    double acc = 0.0, mx = 0.0;
    bool first = true;
    for (int j = 0; j < kN; ++j) {
        const int k = j % kSites;
        for (int a = 0; a < 3; ++a) {
            for (int dm = -1; dm <= 1; ++dm) {
                const int m = k + dm;
                if (m < 0 || m >= kSites) continue;
                const double v = jentry(s, f, a * kSites + m, j);
                acc = first ? v * v : acc + v * v;
                mx = first ? __builtin_fabs(v) : jmax(mx, __builtin_fabs(v));
                first = false;
            }
        }
    }
    if (isfin(acc) && acc >= double(kN * kN) * kDblMin) return __builtin_sqrt(acc);
    if (mx == 0.0 || !isfin(mx)) return mx;
    const int e = exponent_of(mx);
    double t = 0.0;
    first = true;
    for (int j = 0; j < kN; ++j) {
        const int k = j % kSites;
        for (int a = 0; a < 3; ++a) {
            for (int dm = -1; dm <= 1; ++dm) {
                const int m = k + dm;
                if (m < 0 || m >= kSites) continue;
                const double v = ldexp2(jentry(s, f, a * kSites + m, j), -e);
                t = first ? v * v : t + v * v;
                first = false;
            }
        }
    }
    return ldexp2(__builtin_sqrt(t), e);
}

// Jᴴ (sc F) into out (column folds over the nonzero rows).
GS_DEV void jtmatvec(MidShared& s, const Fpu& f, const double* v, double sc, double* out) {
    const int l = lane_id();
    if (l < kN) {
        const int k = l % kSites;
        double acc = 0.0;
        bool first = true;
        for (int a = 0; a < 3; ++a) {
            for (int dm = -1; dm <= 1; ++dm) {
                const int m = k + dm;
                if (m < 0 || m >= kSites) continue;
                const int i = a * kSites + m;
                const double Jik = jentry(s, f, i, l);
                const double vi = v[i] * sc;
                acc = first ? Jik * vi : __builtin_fma(Jik, vi, acc);
                first = false;
            }
        }
        out[l] = acc;
    }
    sync();
}

// grad_ratio_and_norm (newton.jl:1287-1298) at the point of K, with F in s.Fx.
GS_DEV void grad_ratio(MidShared& s, const Fpu& f, double fnorm, double& gr, double& jn) {
    jn = jnorm_fold(s, f);
    if (jn == 0.0 || fnorm == 0.0) {
        gr = 0.0;
        return;
    }
    jtmatvec(s, f, s.Fx, 1.0, s.t3);
    const double gn = fold_rnorm(s.t3);
    if (isfin(gn) || !(isfin(jn) && isfin(fnorm))) {
        gr = gn / jn / fnorm;
        return;
    }
    const int e = exponent_of(fnorm);
    jtmatvec(s, f, s.Fx, ldexp2(1.0, -e), s.t3);
    gr = fold_rnorm(s.t3) / jn / ldexp2(fnorm, -e);
}

// ---- factorisations ------------------------------------------------------------------------------

// The exponent of the largest |Jreg entry| (matrix_exponent, factors.jl:826-843).
GS_DEV int jreg_exponent(const MidShared& s, const Fpu& f, double mu) {
    const int l = lane_id();
    double m = 0.0;
    bool nan_seen = false;
    if (l < kN) {
        const int mm = l % kSites;
        for (int b = 0; b < 3; ++b)
            for (int dk = -1; dk <= 1; ++dk) {
                const int k = mm + dk;
                if (k < 0 || k >= kSites) continue;
                const double v = __builtin_fabs(jreg(s, f, mu, l, b * kSites + k));
                m = __builtin_fmax(m, v);
                nan_seen |= isnan_(v);
            }
    }
    const bool nan_any = __any(nan_seen);
    m = wave_fmax(m);
    if (nan_any || !isfin(m) || m == 0.0) return 0;
    const int e = exponent_of(m);
    return e < -kEmax64 ? -kEmax64 : (e > kEmax64 ? kEmax64 : e);
}

// factor_matrix (factors.jl:802-811) into s.A: Jreg/2^e in FP64 (e < 0 only) or FP32 (any e ≠ 0).
template <bool F32> GS_DEV int build_factor_matrix(MidShared& s, const Fpu& f, double mu) {
    const int e0 = jreg_exponent(s, f, mu);
    const int e = F32 ? e0 : (e0 < 0 ? e0 : 0);
    const int l = lane_id();
    if (l < kN) {
        for (int j = 0; j < kN; ++j) {
            const double v = jreg(s, f, mu, l, j);
            const double w = e != 0 ? ldexp2(v, -e) : v;
            if (F32)
                reinterpret_cast<float*>(s.A)[l + kN * j] = float(w);
            else
                s.A[l + kN * j] = w;
        }
    }
    sync();
    return e;
}

// The pivot of column k over rows k..47: the first largest |a|, or k when |a_kk| is NaN (the
// reference's strict > scan never moves off a NaN and never picks one).
template <class T> GS_DEV int pivot_of(const T* A, int k) {
    const int l = lane_id();
    const double akk = __builtin_fabs(double(A[k + kN * k]));
    double c = -__builtin_inf();
    if (l >= k && l < kN) {
        const double a = __builtin_fabs(double(A[l + kN * k]));
        c = isnan_(a) ? -__builtin_inf() : a;
    }
    const double m = wave_fmax(c);
    if (isnan_(akk)) return k;
    const uint64_t hit = __ballot(l >= k && l < kN && c == m);
    return hit ? __ffsll(static_cast<long long>(hit)) - 1 : k;
}

// Swap rows k and p across every column.
template <class T> GS_DEV void swap_rows(T* A, int k, int p) {
    const int l = lane_id();
    if (p != k && l < kN) {
        const T u = A[k + kN * l];
        A[k + kN * l] = A[p + kN * l];
        A[p + kN * l] = u;
    }
    sync();
}

// Elimination step k on columns k+1 .. jend−1: multipliers a_ik / a_kk, then a_ij − a_ik a_kj.
template <class T> GS_DEV void eliminate(T* A, int k, int jend, int& info) {
    const T akk = A[k + kN * k];
    const int l = lane_id();
    if (akk == T(0)) {
        if (info == 0) info = k + 1;
        return;
    }
    if (l > k && l < kN) {
        const T lik = A[l + kN * k] / akk;
        A[l + kN * k] = lik;
        for (int j = k + 1; j < jend; ++j) A[l + kN * j] = A[l + kN * j] - lik * A[k + kN * j];
    }
    sync();
}

// factorize!! StaticLU (lu.jl:24-64) in FP64, unblocked.
GS_DEV int lu64(MidShared& s) {
    int info = 0;
    for (int k = 0; k < kN; ++k) {
        const int p = pivot_of(s.A, k);
        if (lane_id() == 0) s.piv[k] = p;
        swap_rows(s.A, k, p);
        eliminate(s.A, k, kN, info);
    }
    sync();
    return info;
}

// The same factorisation in FP32, blocked by 16 columns: each panel as the unblocked algorithm
// (its row swaps across all columns), U12 by the panel's unit-lower forward substitution, and the
// trailing update A22 −= L21 U12 on MFMA (v_mfma_f32_16x16x4f32, −L21 as the A operand).
GS_DEV int lu32_blocked(MidShared& s) {
    float* A = reinterpret_cast<float*>(s.A);
    const int l = lane_id();
    int info = 0;
    for (int p0 = 0; p0 < kN; p0 += 16) {
        const int pend = p0 + 16;
        for (int k = p0; k < pend; ++k) {
            const int p = pivot_of(A, k);
            if (l == 0) s.piv[k] = p;
            swap_rows(A, k, p);
            eliminate(A, k, pend, info);
        }
        if (pend >= kN) break;
        // U12: rows p0..pend−1, columns pend..47 (one lane per column)
        for (int k = p0; k < pend; ++k) {
            const int j = pend + l;
            if (j < kN) {
                const float akj = A[k + kN * j];
                for (int i = k + 1; i < pend; ++i) A[i + kN * j] = A[i + kN * j] - A[i + kN * k] * akj;
            }
            sync();
        }
        // A22 −= L21 U12, 16 × 16 blocks, K = 16 in four MFMA steps of 4
        const int nb = (kN - pend) / 16;
        const int rr = l & 15, gg = l >> 4;
        for (int bi = 0; bi < nb; ++bi) {
            for (int bj = 0; bj < nb; ++bj) {
                const int R0 = pend + 16 * bi, C0 = pend + 16 * bj;
                f32x4 acc;
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i] = A[(R0 + 4 * gg + i) + kN * (C0 + rr)];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int kc = p0 + 4 * kk + gg;
                    const float av = -A[(R0 + rr) + kN * kc];
                    const float bv = A[kc + kN * (C0 + rr)];
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
                }
                sync();
#pragma unroll
                for (int i = 0; i < 4; ++i) A[(R0 + 4 * gg + i) + kN * (C0 + rr)] = acc[i];
                sync();
            }
        }
    }
    sync();
    return info;
}

// backsolve (lu.jl:67-92) of the factor in place on x (T = double or float).
template <class T> GS_DEV void backsolve(MidShared& s, const T* A, T* x) {
    const int l = lane_id();
    if (l == 0) {
        for (int k = 0; k < kN; ++k) {
            const int p = s.piv[k];
            if (p != k) {
                const T u = x[k];
                x[k] = x[p];
                x[p] = u;
            }
        }
    }
    sync();
    for (int k = 0; k < kN; ++k) {
        const T xk = x[k];
        if (l > k && l < kN) x[l] = x[l] - A[l + kN * k] * xk;
        sync();
    }
    for (int k = kN - 1; k >= 0; --k) {
        const T xk = x[k] / A[k + kN * k];
        if (l == k) x[k] = xk;
        if (l < k) x[l] = x[l] - A[l + kN * k] * xk;
        sync();
    }
}

// ldiv!! with FP64 factors (factors.jl:931-936): out = J \ r (out ≠ r).
GS_DEV void ldiv64(MidShared& s, int fe, const double* r, double* out) {
    const int k = vec_rhs_exponent(r, fe);
    const int l = lane_id();
    if (l < kN) out[l] = k != 0 ? ldexp2(r[l], -k) : r[l];
    sync();
    backsolve<double>(s, s.A, out);
    const int e = k - fe;
    if (e != 0 && l < kN) out[l] = ldexp2(out[l], e);
    sync();
}

// ldiv!! with FP32 factors (factors.jl:943-949): r/s in FP32, then a·frac·2^(ex−e) in FP64.
GS_DEV void ldiv32(MidShared& s, int fe, const double* r, double* out) {
    const int l = lane_id();
    const double sm = fold_rmax(r);
    if (sm == 0.0) {
        if (l < kN) out[l] = 0.0;
        sync();
        return;
    }
    if (l < kN) s.t32[l] = float(r[l] / sm);
    sync();
    backsolve<float>(s, reinterpret_cast<const float*>(s.A), s.t32);
    const int ex = isfin(sm) ? __builtin_amdgcn_frexp_exp(sm) : 0;
    const double frac = isfin(sm) ? __builtin_amdgcn_frexp_mant(sm) : sm;
    if (l < kN) out[l] = ldexp2(double(s.t32[l]) * frac, ex - fe);
    sync();
}

// ---- the Newton step -----------------------------------------------------------------------------

struct Info {
    uint8_t code;
    double fnorm, stepnorm;
    int32_t failures;
    bool promoted;
};

struct NewtonState {
    double fnorm;      // the ‖F‖ the state carries
    double gradratio;  // diagnostics of the last factorisation
    double jnorm;
    bool promoted;
    int32_t slow;
    double best_fnorm;
};

// globalize!! Static (linefunction.jl:114-131): x, Fx updated on success.
GS_DEV Info globalize_static(MidShared& s, const Fpu& f, const DevParams& q) {
    const int l = lane_id();
    const double dn = fold_rnorm(s.d);
    const double stepnorm = __builtin_fabs(q.alpha) * dn;
    if (!isfin(stepnorm)) return Info{kNonfinite, fold_rnorm(s.Fx), stepnorm, 0, false};
    if (l < kN) s.t1[l] = s.x[l] + q.alpha * s.d[l];
    sync();
    eval_F(s, f, s.t1, s.t2);
    const double fnp = fold_rnorm(s.t2);
    if (!isfin(fnp)) return Info{kNonfinite, fold_rnorm(s.Fx), stepnorm, 0, false};
    if (l < kN) {
        s.x[l] = s.t1[l];
        s.Fx[l] = s.t2[l];
    }
    sync();
    return Info{kSuccess, fnp, stepnorm, 0, false};
}

// refine_passes!! (refinement.jl:211-244) on s.d with the held Jreg; `failed` as the reference's.
// Without Refinement() it leaves d as it is and does not fail.
template <bool F32, bool REFINE>
GS_DEV void refine(MidShared& s, const Fpu& f, const DevParams& q, int fe, bool& failed) {
    failed = false;
    if (!REFINE) return;
    const int l = lane_id();
    const double fnorm = fold_rnorm(s.Fx);
    if (fnorm == 0.0) return;
    if (!isfin(fnorm)) {
        failed = true;
        return;
    }
    jmatvec(s, f, q.mu, s.d, s.t1);
    if (l < kN) s.r[l] = -s.Fx[l] + -s.t1[l];
    sync();
    double rprev = fold_rnorm(s.r);
    if (!isfin(rprev)) {
        failed = true;
        return;
    }
    if (rprev <= q.ref_eta_eff * fnorm) return;
    for (int it = 1; it <= q.ref_maxiter; ++it) {
        if (F32)
            ldiv32(s, fe, s.r, s.t2);
        else
            ldiv64(s, fe, s.r, s.t2);
        if (l < kN) s.t3[l] = s.d[l] + s.t2[l];
        sync();
        jmatvec(s, f, q.mu, s.t3, s.t1);
        if (l < kN) s.r[l] = -s.Fx[l] + -s.t1[l];
        sync();
        const double rcur = fold_rnorm(s.r);
        if (!(rcur < rprev)) {
            failed = true;
            return;
        }
        if (l < kN) s.d[l] = s.t3[l];
        sync();
        rprev = rcur;
        if (rcur <= q.ref_eta_eff * fnorm) return;
    }
    failed = true;
}

// The step's direction d = −(J \ F) into s.d from the factor in s.A.
template <bool F32> GS_DEV void direction(MidShared& s, int fe) {
    if (F32)
        ldiv32(s, fe, s.Fx, s.d);
    else
        ldiv64(s, fe, s.Fx, s.d);
    const int l = lane_id();
    if (l < kN) s.d[l] = -s.d[l];
    sync();
}

// newton_factor (newton.jl:1302-1311) up to the factorisation: fnorm, J at x, its diagnostics.
GS_DEV void factor_diag(MidShared& s, const Fpu& f, NewtonState& st) {
    const double fnorm = isnan_(st.fnorm) ? fold_rnorm(s.Fx) : st.fnorm;
    eval_K(s, f, s.x);
    grad_ratio(s, f, fnorm, st.gradratio, st.jnorm);
}

// newton_step!! plain (newton.jl:1679-1688).
template <bool REFINE> GS_DEV Info plain_step(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st) {
    factor_diag(s, f, st);
    const int fe = build_factor_matrix<false>(s, f, q.mu);
    const int info = lu64(s);
    Info r;
    if (info > 0) {
        r = Info{kSingular, fold_rnorm(s.Fx), 0.0, 0, false};
    } else {
        direction<false>(s, fe);
        bool failed;
        refine<false, REFINE>(s, f, q, fe, failed);
        r = globalize_static(s, f, q);
    }
    st.fnorm = r.fnorm;
    return r;
}

// linear_residual_ratio (newton.jl:1422-1430).
GS_DEV double linear_residual_ratio(MidShared& s, const Fpu& f, const DevParams& q, double fnorm_old) {
    if (fnorm_old == 0.0 || !isfin(fnorm_old)) return 0.0;
    jmatvec(s, f, q.mu, s.d, s.t1);
    const int l = lane_id();
    if (l < kN) s.t2[l] = 1.0 * s.Fx[l] + 1.0 * s.t1[l];
    sync();
    return fold_rnorm(s.t2) / fnorm_old;
}

// mixed_step!! + mixed_newton!! with fresh factors (newton.jl:1604-1666, 1692-1720).
template <bool REFINE> GS_DEV Info mixed_step(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st) {
    const double fnorm_old = isnan_(st.fnorm) ? fold_rnorm(s.Fx) : st.fnorm;
    const bool was_promoted = st.promoted;
    bool ok = true, md_promoted = st.promoted;
    int32_t failures = 0;
    double lin_rel = 0.0;
    if (st.promoted) {
        factor_diag(s, f, st);
        const int fe = build_factor_matrix<false>(s, f, q.mu);
        if (lu64(s) > 0) {
            ok = false;
        } else {
            direction<false>(s, fe);
            bool failed;
            refine<false, REFINE>(s, f, q, fe, failed);
        }
    } else {
        factor_diag(s, f, st);
        const int fe = build_factor_matrix<true>(s, f, q.mu);
        const int info = lu32_blocked(s);
        bool ref_failed = false;
        if (info > 0) {
            ref_failed = true;
        } else {
            direction<true>(s, fe);
            refine<true, REFINE>(s, f, q, fe, ref_failed);
        }
        if (ref_failed || info > 0) {
            failures += 1;
            // promote_factor!! (newton.jl:1332-1339): Jreg in FP64
            const int fe64 = build_factor_matrix<false>(s, f, q.mu);
            md_promoted = true;
            st.promoted = true;
            if (lu64(s) > 0) {
                ok = false;
            } else {
                direction<false>(s, fe64);
            }
        } else {
            lin_rel = linear_residual_ratio(s, f, q, fnorm_old);
        }
    }
    if (!ok) {
        Info r{kSingular, fold_rnorm(s.Fx), 0.0, failures, st.promoted};
        st.fnorm = r.fnorm;
        return r;
    }
    const bool promoted_step = md_promoted && !was_promoted;
    Info r = globalize_static(s, f, q);
    r.failures += failures;
    if (promoted_step) {
        r.promoted = true;
        st.promoted = true;
        st.slow = 0;
    } else if (!was_promoted) {
        // read_mixed_stall (newton.jl:1437-1451)
        const double eta_lin = 1e-4;  // max(√eps, 1e-4)
        if (r.fnorm > 0.5 * fnorm_old && lin_rel > eta_lin) {
            const int32_t slow = st.slow + 1;
            if (slow >= 2) {
                st.promoted = true;
                st.slow = 0;
                r.failures += 1;
                r.promoted = true;
            } else {
                st.promoted = false;
                st.slow = slow;
            }
        } else {
            st.promoted = false;
            st.slow = 0;
        }
    }
    st.fnorm = r.fnorm;
    return r;
}

// ---- the solve ----------------------------------------------------------------------------------

GS_DEV void store_status_mid(gs_status_f64* dst, uint8_t code, int32_t iterations, double fnorm, double stepnorm,
                             int32_t failures, bool promoted, double gradratio, double rate, bool contracted) {
    uint64_t* w = reinterpret_cast<uint64_t*>(dst);
    w[0] = uint64_t(code) | (uint64_t(uint32_t(iterations)) << 32);
    w[1] = __builtin_bit_cast(uint64_t, fnorm);
    w[2] = __builtin_bit_cast(uint64_t, stepnorm);
    w[3] = uint64_t(uint32_t(failures)) | (uint64_t(promoted ? 1 : 0) << 32);
    w[4] = uint64_t(uint32_t(-1));
    w[5] = __builtin_bit_cast(uint64_t, dnan());
    w[6] = __builtin_bit_cast(uint64_t, gradratio);
    w[7] = __builtin_bit_cast(uint64_t, rate);
    w[8] = uint64_t(contracted ? 1 : 0);
}

template <bool MIXED, bool REFINE> __global__ __launch_bounds__(64) void fpu_gl3_kernel(const MidLaunch L) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MidShared& s = *reinterpret_cast<MidShared*>(lds_raw);
    const int l = lane_id();
    const int64_t i = blockIdx.x;
    const int64_t idx = L.order ? L.order[i] - 1 : i;
    const DevParams& q = L.q;
    const Fpu f = fpu_of(L.shared);
    {
        const double* P = reinterpret_cast<const double*>(L.P + (L.p_stride ? idx * L.p_stride : 0));
        if (l < 32) s.qp[l] = P[l];
        if (l < kN) s.x[l] = L.X[idx * kN + l];
    }
    sync();
    // solve!! (solver.jl:277-378)
    eval_F(s, f, s.x, s.Fx);
    const double fnorm0 = fold_rnorm(s.Fx);
    NewtonState st{fnorm0, dnan(), 0.0, false, 0, fnorm0};
    if (q.best && l < kN) s.bx[l] = s.x[l];
    uint8_t code = kSuccess;
    int32_t it = 0, failures = 0;
    double fnorm = fnorm0, stepnorm = 0.0, rate = dnan();
    bool promoted = false, contracted = false;
    bool done = false;
    if (!isfin(fnorm0)) {
        code = kNonfinite;
        done = true;
    }
    double ftest0 = fnorm0;
    if (!done) {
        ftest0 = q.norm_max ? fold_rmax(s.Fx) : fnorm0;
        if (0 >= q.min_iterations && ftest0 <= jmax(q.f_abstol, q.f_reltol * ftest0)) done = true;
        else if (q.max_iterations == 0) {
            code = kMaxIters;
            done = true;
        }
    }
    int32_t stalls = 0;
    while (!done) {
        const double rate_k = rate;
        const Info info = MIXED ? mixed_step<REFINE>(s, f, q, st) : plain_step<REFINE>(s, f, q, st);
        if (q.best && info.fnorm < st.best_fnorm) {
            if (l < kN) s.bx[l] = s.x[l];
            st.best_fnorm = info.fnorm;
            sync();
        }
        // record (status.jl:245-262)
        rate = stepnorm > 0.0 ? info.stepnorm / stepnorm : dnan();
        code = info.code;
        it += 1;
        fnorm = info.fnorm;
        stepnorm = info.stepnorm;
        failures += info.failures;
        promoted = promoted || info.promoted;
        if (code == kSingular || code == kNonfinite || code == kMaxIters) break;
        const double ftest = q.norm_max ? fold_rmax(s.Fx) : fnorm;
        if (it >= q.min_iterations && ftest <= jmax(q.f_abstol, q.f_reltol * ftest0)) {
            code = kSuccess;
            break;
        }
        const double xnorm = fold_rnorm(s.x);
        const double xtol = jmax(q.x_abstol, q.x_reltol * xnorm);
        const bool past_min = it >= q.min_iterations;
        bool ctr = false;
        if (q.rate_test && past_min) {
            const double th = jmax(rate, rate_k);
            ctr = stepnorm > 0.0 && th < 1.0 && th / (1.0 - th) * stepnorm <= q.kappa * xtol;
        }
        const bool step_conv = past_min && it > 0 && stepnorm > 0.0 && stepnorm <= xtol;
        const bool zero_step = code == kSuccess && info.stepnorm == 0.0 && past_min;
        const bool rstop = q.roundoff && past_min && rate >= 1.0 && rate_k < 1.0;
        if (ctr || step_conv || zero_step || rstop) {
            code = ctr ? kSuccess : kStalled;
            contracted = ctr;
            break;
        }
        if (code == kStalled) {
            stalls += 1;
            if (stalls >= q.max_stalls) break;
        } else {
            stalls = 0;
        }
        if (it >= q.max_iterations) {
            code = kMaxIters;
            break;
        }
    }
    // finish!! (newton.jl:1165-1231): best restore
    if (q.best && code != kSuccess && isfin(st.best_fnorm) && st.best_fnorm < fnorm) {
        if (l < kN) s.x[l] = s.bx[l];
        fnorm = st.best_fnorm;
        sync();
    }
    if (l < kN) L.X[idx * kN + l] = s.x[l];
    if (l == 0) {
        if (L.S) store_status_mid(L.S + idx, code, it, fnorm, stepnorm, failures, promoted, st.gradratio, rate, contracted);
        if (L.codes) L.codes[idx] = code;
    }
}

}  // namespace

template <bool MIXED, bool REFINE> int launch_variant(const MidLaunch& L, hipStream_t st) {
    const size_t lds = sizeof(MidShared);
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&fpu_gl3_kernel<MIXED, REFINE>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
            return set_error(GS_EHIP, "hipFuncSetAttribute failed");
        attr_set = true;
    }
    hipLaunchKernelGGL((fpu_gl3_kernel<MIXED, REFINE>), dim3(unsigned(L.nsys)), dim3(64), lds, st, L);
    return GS_OK;
}

int launch_fpu_gl3(const MidLaunch& L, hipStream_t st) {
    if (L.nsys > 0x7fffffffLL) return set_error(GS_EINVAL, "a wave-per-system batch holds at most 2^31-1 systems");
    if (L.mixed) return L.refine ? launch_variant<true, true>(L, st) : launch_variant<true, false>(L, st);
    return L.refine ? launch_variant<false, true>(L, st) : launch_variant<false, false>(L, st);
}

}  // namespace gs
