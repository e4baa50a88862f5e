// libgsolve_hip, mid-size systems: the wave-per-system solve of n = 48 (the device code that
// mid_solver.hip, with the plain launches, and mid_reuse.hip, with the reuse = k launches, share).
//
// The FPU-β Gauss–Legendre s = 3 stage system (n = 48, GS_FAM_FPU_GL3) under Newton with the Static
// or the Backtracking globalization, with or without reuse = k, in the working precision (static LU
// in FP64) or mixed (LUFactorization(Float32) with Refinement and promotion to FP64): the
// reference's R1 loop solve!! (src/nonlinear/solver.jl:277-378) with newton_step!!
// (src/nonlinear/newton.jl:1604-1874) per system, as the lane-per-system kernels run it for n ≤ 8.
//
// Lane i < 48 owns entry i of every vector (in LDS) and row i of the factor (in registers: 96 VGPRs
// in FP64, 48 in FP32); the scalar state of the solve is uniform across the wave. The factorisation
// is the reference's unblocked LU in the factor's precision — first-largest pivot, a/p multipliers,
// rank-1 updates without FMA — so both the FP64 and the FP32 factor match the oracle bit for bit:
//   - the pivot search is a DPP max over the wave, then a ballot of the lanes that hold it;
//   - rows never move: each lane keeps the logical row index `pos` of the row it holds, and a row
//     swap exchanges two indices (ties go to the smallest logical index, as the reference's scan);
//   - the pivot row's entries reach the other lanes by v_readlane, one broadcast per entry, which
//     every lane below the pivot uses in one multiply and one subtract.
// In this layout a step of the rank-1 update costs one instruction per column for all 48 rows at
// once, so a blocked form with the trailing update on MFMA saves no VALU work and adds LDS staging
// (measured: the LDS-resident blocked FP32 LU took 128k cycles per factorisation).
//
// Sums that the reference folds left to right (norms, the matrix-vector products, ‖J‖_F) are folded
// in that order: the norms by every lane over the LDS vector (broadcast reads), the products of the
// tridiagonal-block Jacobian by the lane of their row or column over its nonzero entries only (a
// structural zero adds an exact zero to the fold). The Jacobian itself is never stored: its entries
// are recomputed from the bond stiffnesses K.
#pragma once

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
    double x[kN], Fx[kN], d[kN], r[kN], t1[kN], t2[kN], t3[kN], t4[kN], bx[kN];
    double jx[kN];      // the point of the last factorisation's Jacobian
    double jsq[kN * 9]; // the squares of ‖J‖_F's nonzero entries, column by column
    double qp[32];      // q, p of the system
    double Qv[kN];      // stage positions
    double g[kN];       // stage forces
    double K[3 * 17];   // bond stiffnesses at the Jacobian's point
};

GS_DEV void sync() { __syncthreads(); }

// LDS written by some lanes of the wave, then read by others (the workgroup is the one wave).
GS_DEV void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

GS_DEV int lane_id() { return int(threadIdx.x); }

// ---- wave primitives (called with all 64 lanes active) ---------------------------------------------

GS_DEV int readlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
GS_DEV float readlane(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
GS_DEV double readlane(double v, int lane) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), lane));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), lane));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// The value a DPP move (CTRL) brings into this lane; lanes outside the row mask ROWS keep their own.
template <int CTRL, int ROWS> GS_DEV float dpp_move(float v) {
    const int b = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS> GS_DEV double dpp_move(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = int(uint32_t(b)), hi = int(uint32_t(b >> 32));
    const uint32_t mlo = uint32_t(__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xF, false));
    const uint32_t mhi = uint32_t(__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xF, false));
    return __builtin_bit_cast(double, (uint64_t(mhi) << 32) | mlo);
}

// The maximum over the wave of a value that is nowhere NaN, in every lane: quad swaps, half-row and
// row mirrors give each row of 16 its maximum, the row_bcast:15 / row_bcast:31 steps carry it into
// lane 63.
template <class T> GS_DEV T wave_max(T v) {
    v = tfmax(v, dpp_move<0xB1, 0xF>(v));   // quad_perm [1,0,3,2]
    v = tfmax(v, dpp_move<0x4E, 0xF>(v));   // quad_perm [2,3,0,1]
    v = tfmax(v, dpp_move<0x141, 0xF>(v));  // row_half_mirror
    v = tfmax(v, dpp_move<0x140, 0xF>(v));  // row_mirror
    v = tfmax(v, dpp_move<0x142, 0xA>(v));  // row_bcast:15 into rows 1 and 3
    v = tfmax(v, dpp_move<0x143, 0xC>(v));  // row_bcast:31 into rows 2 and 3
    return readlane(v, 63);
}

// The signed maximum of a 32-bit integer over the wave, in every lane (the same DPP pattern on
// v_max_i32).
template <int CTRL, int ROWS> GS_DEV int dpp_move(int v) {
    // lanes outside ROWS receive INT_MIN, the identity of the max (so the move folds into v_max_i32_dpp)
    return __builtin_amdgcn_update_dpp(int(0x80000000u), v, CTRL, ROWS, 0xF, false);
}
GS_DEV int imax(int a, int b) { return a > b ? a : b; }
GS_DEV int wave_max_i32(int v) {
    v = imax(v, dpp_move<0xB1, 0xF>(v));
    v = imax(v, dpp_move<0x4E, 0xF>(v));
    v = imax(v, dpp_move<0x141, 0xF>(v));
    v = imax(v, dpp_move<0x140, 0xF>(v));
    v = imax(v, dpp_move<0x142, 0xA>(v));
    v = imax(v, dpp_move<0x143, 0xC>(v));
    return readlane(v, 63);
}

GS_DEV int lowest_lane(uint64_t mask) { return __builtin_ctzll(mask); }

// The lanes holding the largest of c over the lanes with `cand`, where c is +0, a positive float,
// +inf or −inf (a non-candidate or a NaN): as IEEE bit patterns these order as signed integers, so
// the maximum is an integer reduction — in FP64 over the high words, then over the low words of the
// lanes that share the largest high word. (Each reduction runs before any condition that depends on
// the lane, so that every lane takes part.)
GS_DEV uint64_t lanes_of_max(float c, bool cand) {
    const int key = __builtin_bit_cast(int, c);
    const int top = wave_max_i32(key);
    return __ballot(cand && key == top);
}
GS_DEV uint64_t lanes_of_max(double c, bool cand) {
    const uint64_t b = __builtin_bit_cast(uint64_t, c);
    const int hi = int(uint32_t(b >> 32));
    const int mhi = wave_max_i32(hi);
    const bool top = cand && hi == mhi;
    uint64_t hit = __ballot(top);
    if (hit & (hit - 1)) {
        const int lo = top ? int(uint32_t(b) ^ 0x80000000u) : int(0x80000000u);
        const int mlo = wave_max_i32(lo);
        hit = __ballot(top && lo == mlo);
    }
    return hit;
}

// Materialises v in its VGPRs here: an empty asm that neither reads memory nor emits an instruction.
// It stops the compiler from sinking a per-lane select to the select's far use, which would keep
// both inputs and the lane mask of every LU step live across the whole factorisation.
template <class T> GS_DEV void keep(T& v) { asm volatile("" : "+v"(v)); }

// Left folds over an LDS vector of 48, evaluated by every lane (uniform results). The loops unroll
// by 8 only: a full unroll would hoist all 48 broadcast reads into 96 VGPRs beside the factor rows.
GS_DEV double fold_norm2(const double* v) {
    double s = v[0] * v[0];
#pragma unroll 8
    for (int i = 1; i < kN; ++i) s = s + v[i] * v[i];
    return s;
}

GS_DEV double fold_rmax(const double* v) {
    double m = __builtin_fabs(v[0]);
#pragma unroll 8
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
#pragma unroll 8
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
    const bool nan_any = __ballot(l < kN && isnan_(a)) != 0;
    const double m = wave_max(isnan_(a) ? 0.0 : a);
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

// ‖J‖_F: the reference's left fold over the entries in column-major order (rnorm, reductions.jl:
// 41-57). Lane j forms the nine entries of column j at rows a·16 + k−1, k, k+1 (k = j mod 16; zero
// where the chain ends, and a zero square adds nothing to the nonnegative sum), then the fold takes
// their squares column by column through v_readlane.
GS_DEV double jnorm_fold(MidShared& s, const Fpu& f) {
    const int l = lane_id();
    const int j = l < kN ? l : 0;
    const int k = j % kSites;
    double v[9], sq[9];
    double mx = 0.0;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int m = k + t % 3 - 1;
        v[t] = (l < kN && m >= 0 && m < kSites) ? jentry(s, f, (t / 3) * kSites + m, j) : 0.0;
        sq[t] = v[t] * v[t];
        mx = jmax(mx, __builtin_fabs(v[t]));
        if (l < kN) s.jsq[9 * l + t] = sq[t];
    }
    wsync();
    // the fold over columns 0..47, entries t = 0..8 of each: the LDS array in order, read by every
    // lane as 16-byte broadcasts
    double acc = 0.0;
    const double2* q = reinterpret_cast<const double2*>(s.jsq);
#pragma unroll 8
    for (int c = 0; c < kN * 9 / 2; ++c) {
        const double2 w = q[c];
        acc = acc + w.x;
        acc = acc + w.y;
    }
    if (isfin(acc) && acc >= double(kN * kN) * kDblMin) return __builtin_sqrt(acc);
    // the scaled path: the NaN-propagating maximum of |J_ij|, then the fold of (J_ij/2^e)²
    const bool nan_any = __ballot(isnan_(mx)) != 0;
    mx = nan_any ? dnan() : wave_max(mx);
    if (mx == 0.0 || !isfin(mx)) return mx;
    const int e = exponent_of(mx);
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const double w = ldexp2(v[t], -e);
        sq[t] = w * w;
    }
    double acc2 = 0.0;
    for (int c = 0; c < kN; ++c) {
#pragma unroll
        for (int t = 0; t < 9; ++t) acc2 = acc2 + readlane(sq[t], c);
    }
    return ldexp2(__builtin_sqrt(acc2), e);
}

// Jᴴ (sc F) into out (column folds over the nonzero rows).
GS_DEV void jtmatvec(MidShared& s, const Fpu& f, double mu, const double* v, double sc, double* out) {
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
                const double Jik = jreg(s, f, mu, i, l);
                const double vi = v[i] * sc;
                acc = first ? Jik * vi : __builtin_fma(Jik, vi, acc);
                first = false;
            }
        }
        out[l] = acc;
    }
    sync();
}

// grad_ratio_and_norm (newton.jl:1287-1298) at the point of K, with F in Fv (an LDS vector ≠ s.t3).
GS_DEV void grad_ratio(MidShared& s, const Fpu& f, const double* Fv, double fnorm, double& gr, double& jn) {
    jn = jnorm_fold(s, f);
    if (jn == 0.0 || fnorm == 0.0) {
        gr = 0.0;
        return;
    }
    jtmatvec(s, f, 0.0, Fv, 1.0, s.t3);
    const double gn = fold_rnorm(s.t3);
    if (isfin(gn) || !(isfin(jn) && isfin(fnorm))) {
        gr = gn / jn / fnorm;
        return;
    }
    const int e = exponent_of(fnorm);
    jtmatvec(s, f, 0.0, Fv, ldexp2(1.0, -e), s.t3);
    gr = fold_rnorm(s.t3) / jn / ldexp2(fnorm, -e);
}

// ---- factorisations ------------------------------------------------------------------------------

// factor_matrix (factors.jl:802-811) as rows: lane l < 48 receives row l of Jreg/2^e in T, with e
// the exponent of the largest |Jreg entry| (matrix_exponent, factors.jl:826-843) applied when ≠ 0
// in FP32 and only when < 0 in FP64; returns e. Row i = a·16 + m has its nonzeros at columns
// b·16 + m + d (d = −1, 0, 1, within the chain): the lane forms those nine entries (and the
// structural zero 0 − c_ab·0 of each block) once, then places them with three compares per column.
// `tame` reports, uniformly, that every entry is finite, no entry is −0 and every |entry| <
// 2^(emax−48): then no step of the LU meets a non-finite value (growth ≤ 2^47 under partial
// pivoting, multipliers ≤ 1) nor makes a −0, and a zero multiplier leaves a row exactly as it is.
template <class T> GS_DEV int factor_rows(const MidShared& s, const Fpu& f, double mu, T (&a)[kN], bool& tame) {
    constexpr bool kF32 = sizeof(T) == 4;
    const int l = lane_id();
    const bool live = l < kN;
    const int i = live ? l : kN - 1;
    const int ab = i / kSites, m = i % kSites;
    double band[3][3], zero[3];
    double mx = 0.0;
    bool nan_seen = false;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const double km = s.K[b * 17 + m], kp = s.K[b * 17 + m + 1], cf = f.coef(ab, b);
        zero[b] = 0.0 - cf * 0.0;
        nan_seen |= isnan_(zero[b]);
#pragma unroll
        for (int d = -1; d <= 1; ++d) {
            const double H = d == 0 ? -(km + kp) : (d < 0 ? km : kp);
            const bool diag = d == 0 && b == ab;
            double v = (diag ? 1.0 : 0.0) - cf * H;
            if (diag && mu != 0.0) v = v + mu;
            band[b][d + 1] = v;
            const bool in_row = live && m + d >= 0 && m + d < kSites;
            if (in_row) {
                mx = __builtin_fmax(mx, __builtin_fabs(v));
                nan_seen |= isnan_(v);
            }
        }
    }
    const bool nan_any = __ballot(live && nan_seen) != 0;
    const double wmx = wave_max(isnan_(mx) ? 0.0 : mx);
    int e0 = 0;
    if (!nan_any && isfin(wmx) && wmx != 0.0) {
        const int ex = exponent_of(wmx);
        e0 = ex < -kEmax64 ? -kEmax64 : (ex > kEmax64 ? kEmax64 : ex);
    }
    const int e = kF32 ? e0 : (e0 < 0 ? e0 : 0);
    const T big = ldexp2(T(1), Lim<T>::emax - kN);
    bool wild = false;
    T tb[3][3], tz[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        tz[b] = T(e != 0 ? ldexp2(zero[b], -e) : zero[b]);
        wild |= !(tabs(tz[b]) < big) || (tz[b] == T(0) && __builtin_signbit(tz[b]));
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            tb[b][d] = T(e != 0 ? ldexp2(band[b][d], -e) : band[b][d]);
            const bool in_row = m + d - 1 >= 0 && m + d - 1 < kSites;
            wild |= in_row && (!(tabs(tb[b][d]) < big) || (tb[b][d] == T(0) && __builtin_signbit(tb[b][d])));
        }
    }
#pragma unroll
    for (int j = 0; j < kN; ++j) {
        const int b = j / kSites, k = j % kSites;
        a[j] = k + 1 == m ? tb[b][0] : (k == m ? tb[b][1] : (k == m + 1 ? tb[b][2] : tz[b]));
    }
    tame = __ballot(live && wild) == 0;
    return e;
}

// factorize!! StaticLU (lu.jl:24-64) on register rows; returns info (the first zero pivot, 1-based,
// or 0). `pos` receives the logical row each lane's row ends in (−1 in lanes 48..63). Step k: the
// pivot is the first largest |a_ik| over the logical rows i ≥ k, or row k itself when |a_kk| is NaN
// (the reference's strict > scan never leaves a NaN and never picks one); the swap of rows k and p
// exchanges two `pos` values; then every lane below the pivot forms its multiplier a_ik/a_kk and
// subtracts multiplier × pivot-row entry (no FMA) from each entry right of column k.
// TAME: the rows satisfy factor_rows' `tame`, and the other lanes run the same update with a zero
// multiplier (an exact no-op there), which keeps every entry in its register; otherwise each entry
// is selected per lane.
template <class T, bool TAME> GS_DEV int lu_rows(T (&a)[kN], int& pos) {
    const int l = lane_id();
    const bool live = l < kN;
    pos = live ? l : -1;
    int info = 0;
#pragma unroll
    for (int k = 0; k < kN; ++k) {
        const T ak = a[k];
        const bool cand = live && pos >= k;
        const T mag = tabs(ak);
        const T c = (cand && !isnan_(mag)) ? mag : -T(__builtin_inf());
        const uint64_t hit = lanes_of_max(c, cand);
        const int owner = lowest_lane(__ballot(live && pos == k));
        int piv = owner;
        if (!isnan_(readlane(ak, owner))) {
            piv = lowest_lane(hit);
            if (hit & (hit - 1)) {
                // a tie: the smallest logical row among the lanes holding the maximum
                int best = readlane(pos, piv);
                for (uint64_t h = hit & (hit - 1); h; h &= h - 1) {
                    const int ln = lowest_lane(h);
                    const int pl = readlane(pos, ln);
                    if (pl < best) {
                        best = pl;
                        piv = ln;
                    }
                }
            }
        }
        const int pp = readlane(pos, piv);
        if (l == owner) pos = pp;
        if (l == piv) pos = k;
        // a zero pivot records info and eliminates nothing (no branch: the update runs with no row
        // below, so that no entry changes register)
        const T ukk = readlane(ak, piv);
        const bool zero = ukk == T(0);
        if (zero && info == 0) info = k + 1;
        const bool below = live && pos > k && !zero;
        const T lik = ak / ukk;
        a[k] = below ? lik : ak;
        keep(a[k]);
        if (TAME) {
            const T mk = below ? lik : T(0);
            // the pivot row's entries reach the other lanes by v_readlane (into SGPRs); a sched
            // barrier every eight columns keeps only a few broadcasts in flight
#pragma unroll
            for (int j = k + 1; j < kN; ++j) {
                a[j] = a[j] - mk * readlane(a[j], piv);
                if ((j - k) % 8 == 0) __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int j = k + 1; j < kN; ++j) {
                const T t = a[j] - lik * readlane(a[j], piv);
                a[j] = below ? t : a[j];
                keep(a[j]);
            }
        }
    }
    return info;
}

// backsolve (lu.jl:67-92) on the register rows: y is entry l of the right-hand side in lane l (the
// row permutation is the `pos` map); returns, in each lane, entry pos of the solution. Forward over
// the unit lower factor, then backward over U, each step broadcasting the finished entry.
template <class T> GS_DEV T solve_rows(const T (&a)[kN], int pos, T y) {
    const int l = lane_id();
    const bool live = l < kN;
#pragma unroll
    for (int k = 0; k < kN - 1; ++k) {
        const T yk = readlane(y, lowest_lane(__ballot(live && pos == k)));
        const T t = y - a[k] * yk;
        y = live && pos > k ? t : y;
    }
#pragma unroll
    for (int k = kN - 1; k >= 0; --k) {
        const int o = lowest_lane(__ballot(live && pos == k));
        const T quo = y / a[k];  // the owner's quotient (the other lanes' are not used)
        const T xk = readlane(quo, o);
        const T t = y - a[k] * xk;
        y = l == o ? xk : (live && pos < k ? t : y);
    }
    return y;
}

// ldiv!! with FP64 factors (factors.jl:931-936): out = J \ r (out ≠ r).
GS_DEV void ldiv_rows(MidShared& s, const double (&a)[kN], int pos, int fe, const double* r, double* out) {
    const int k = vec_rhs_exponent(r, fe);
    const int l = lane_id();
    double y = 0.0;
    if (l < kN) y = k != 0 ? ldexp2(r[l], -k) : r[l];
    y = solve_rows<double>(a, pos, y);
    const int e = k - fe;
    if (l < kN) out[pos] = e != 0 ? ldexp2(y, e) : y;
    sync();
}

// ldiv!! with FP32 factors (factors.jl:943-949): r/s in FP32, then a·frac·2^(ex−e) in FP64.
GS_DEV void ldiv_rows(MidShared& s, const float (&a)[kN], int pos, int fe, const double* r, double* out) {
    const int l = lane_id();
    const double sm = fold_rmax(r);
    if (sm == 0.0) {
        if (l < kN) out[l] = 0.0;
        sync();
        return;
    }
    float y = 0.0f;
    if (l < kN) y = float(r[l] / sm);
    y = solve_rows<float>(a, pos, y);
    const int ex = isfin(sm) ? __builtin_amdgcn_frexp_exp(sm) : 0;
    const double frac = isfin(sm) ? __builtin_amdgcn_frexp_mant(sm) : sm;
    if (l < kN) out[pos] = ldexp2(double(y) * frac, ex - fe);
    sync();
}

// ---- the Newton step -----------------------------------------------------------------------------

struct Info {
    uint8_t code;
    double fnorm, stepnorm;
    int32_t failures;
    bool promoted;
    bool shortened = false;  // a line search took less than the full step
    bool limited = false;    // a trust radius, not the Newton step, set the step's norm
};

struct NewtonState {
    double fnorm;      // the ‖F‖ the state carries
    double gradratio;  // diagnostics of the last factorisation, formed at exit
    double jnorm;
    bool promoted;
    int32_t slow;
    double best_fnorm;
    double jfnorm;     // the ‖F‖ the last factorisation's diagnostics read (its point is s.jx)
    bool have_j;
    double merit;      // norm2(Fx) under Backtracking (the line search's φ), NaN under Static
    double radius;     // DogLeg: the trust radius (TrustRadius, newton.jl:637-640), 0 until set
};

// residual_norm (newton.jl:1274-1278): the state's ‖F‖ under Static; under Backtracking √merit
// where the merit is finite and at least n·floatmin, else rnorm(F).
template <int GLOB> GS_DEV double state_fnorm(MidShared& s, const NewtonState& st) {
    if (GLOB == kBacktracking) {
        if (isfin(st.merit) && st.merit >= double(kN) * kDblMin) return __builtin_sqrt(st.merit);
        return fold_rnorm(s.Fx);
    }
    return isnan_(st.fnorm) ? fold_rnorm(s.Fx) : st.fnorm;
}

// rdot (reductions.jl:77-85) of two LDS vectors: a left fold of products, no FMA.
GS_DEV double fold_dot(const double* a, const double* b) {
    double t = a[0] * b[0];
#pragma unroll 8
    for (int i = 1; i < kN; ++i) t = t + a[i] * b[i];
    return t;
}

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

// globalize_merit!! for Backtracking (linefunction.jl:159-199 → linesearch.jl:154-171 →
// backtracking.jl:354-390), as the lane-per-system step_backtracking: `exact` takes ExactStep's slope
// −2φ₀, otherwise MeasuredSlope's φ′(0) = 2⟨F, J d⟩ at x + 0·d. Trials in s.t1 (point) and s.t2
// (residual); x, Fx and the merit updated where a step is taken.
GS_DEV Info globalize_backtracking(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st, bool exact) {
    const int l = lane_id();
    const double phi0 = st.merit;
    const double dn = fold_rnorm(s.d);
    if (!isfin(dn)) return Info{kNonfinite, fold_rnorm(s.Fx), dn, 0, false};
    double d0;
    if (exact) {
        d0 = -2.0 * phi0;
    } else {
        if (l < kN) s.t1[l] = s.x[l] + 0.0 * s.d[l];
        sync();
        eval_F(s, f, s.t1, s.t2);
        eval_K(s, f, s.t1);
        jmatvec(s, f, 0.0, s.d, s.t3);
        d0 = 2.0 * fold_dot(s.t2, s.t3);
    }
    if (!(isfin(phi0) && isfin(d0) && d0 < 0.0)) {
        const uint8_t c = !(isfin(phi0) && isfin(d0)) ? kNonfinite : (d0 > 0.0 ? kLsFailed : kStalled);
        // a LINESEARCH_FAILED anchor keeps x with φ = NaN: the failure takes no step
        if (c == kLsFailed) return Info{kLsFailed, fold_rnorm(s.Fx), 0.0, 1, false};
        return Info{c, fold_rnorm(s.Fx), 0.0, 0, false};
    }
    const double tau = __builtin_fmax(4.0 * kDblEps * __builtin_fabs(phi0), 4.0 * kDblTiny);
    const double amin = jmax(tau / __builtin_fabs(d0), kDblMin);
    double alpha = __builtin_fmin(__builtin_fmax(1.0, amin), __builtin_inf());
    double aa = alpha, pa = dnan(), ap = dnan(), pp = dnan();
    int frozen = 0;
    uint8_t code = kLsFailed;
    const double sqrteps = Lim<double>::sqrteps;
    for (int trips = 0; trips < q.bt_maxiter; ++trips) {
        aa = alpha;
        if (l < kN) s.t1[l] = s.x[l] + alpha * s.d[l];
        sync();
        eval_F(s, f, s.t1, s.t2);
        pa = fold_norm2(s.t2);
        frozen = (pa == phi0 && alpha <= sqrteps) ? frozen + 1 : 0;
        if (pa - phi0 <= jmin(0.0, q.c1 * alpha * d0 + tau)) {
            code = (isfin(pa) && pa - phi0 < -tau) ? kSuccess
                                                   : (__builtin_fabs(pa - phi0) <= tau ? kStalled : kLsFailed);
            break;
        } else if (frozen >= 2) {
            code = kStalled;
            break;
        } else if (alpha <= amin) {
            code = __builtin_fabs(pa - phi0) <= tau ? kStalled : kLsFailed;
            break;
        }
        const double an = isfin(pa) ? backtrack_next<double>(phi0, d0, alpha, pa, ap, pp, q.p)
                                    : jmin(0.1 * alpha, __builtin_sqrt(alpha) * __builtin_sqrt(amin));
        ap = alpha;
        pp = pa;
        alpha = jmax(an, amin);
    }
    const bool take = code == kSuccess || (code == kLsFailed && isfin(pa) && pa < phi0);
    if (take) {
        // the last trial is x + aa d and its residual
        if (l < kN) {
            s.x[l] = s.t1[l];
            s.Fx[l] = s.t2[l];
        }
        sync();
        st.merit = pa;
        Info r{code, __builtin_sqrt(pa), aa * dn, code == kLsFailed ? 1 : 0, false};
        r.shortened = aa < 1.0;
        return r;
    }
    if (code == kLsFailed) return Info{kLsFailed, fold_rnorm(s.Fx), 0.0, 1, false};
    return Info{code == kStalled ? kStalled : kNonfinite, fold_rnorm(s.Fx), 0.0, 0, false};
}

// DogLeg (dogleg.jl:116-206, newton.jl:1725-1780): the radius clamps of the reference.
GS_DEV double dl_clamp(double v) { return v > 1e10 ? 1e10 : (v < 1e-6 ? 1e-6 : v); }
GS_DEV double dl_shrink(double v) { return jmax(0.25 * v, 1e-6); }
GS_DEV double dl_grow(double v) { return jmin(4.0 * v, 1e10); }

// The trial loop of dogleg_step, as the lane-per-system dogleg_trial, on the model matrix Jreg at the
// factorisation's point (the stiffnesses in s.K) with the Newton step s.d (read where `newton`): the
// Newton step inside the radius, else the Cauchy point cut to the radius, else the point of norm δ
// on the segment between them; a trial is accepted when pred > 0 and ared ≥ 1e-4 pred (unsquared
// norms), a rejection shrinks δ to max(δ/4, 1e-6), and a rejection at 1e-6 fails. ĝ in s.r, the
// trial step in s.t3, its point in s.t2 and residual in s.t4; x, Fx and the radius updated.
GS_DEV Info dogleg_trial(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st, bool newton) {
    const int l = lane_id();
    const double lo = 1e-6;
    const double fnorm = fold_rnorm(s.Fx);
    if (fnorm == 0.0) return Info{kSuccess, fnorm, 0.0, 0, false};
    const double nN = newton ? fold_rnorm(s.d) : __builtin_inf();
    // the Cauchy direction from the unit residual: g̃ = Jregᵀ (F/‖F‖), ĝ = g̃/‖g̃‖
    const double finv = 1.0 / fnorm;
    jtmatvec(s, f, q.mu, s.Fx, finv, s.r);
    const double gnorm = fold_rnorm(s.r);
    const double ginv = 1.0 / gnorm;
    if (l < kN) s.r[l] = ginv * s.r[l];
    sync();
    jmatvec(s, f, q.mu, s.r, s.t1);
    const double jnorm = fold_rnorm(s.t1);
    const bool cauchy = gnorm != 0.0 && isfin(gnorm) && jnorm != 0.0 && isfin(jnorm);
    const double nC = fnorm * ((gnorm / jnorm) / jnorm);
    double delta = st.radius;
    if (delta == 0.0 && (newton || cauchy)) {
        const double n0 = newton ? nN : nC;
        delta = n0 < lo ? 2.0 * lo : dl_clamp(n0);
    }
    while (true) {
        bool boundary;
        const bool inside = newton && nN <= delta;
        if (inside) {
            if (l < kN) s.t3[l] = s.d[l];
            boundary = nN >= delta;
        } else if (!cauchy) {
            st.radius = delta;
            return Info{kStalled, fnorm, 0.0, 0, false};
        } else if (!newton || nC >= delta) {
            const double a = -jmin(nC, delta);
            if (l < kN) s.t3[l] = a * s.r[l];
            boundary = nC >= delta;
        } else {
            // the root t ≥ 0 of ‖a + t b̂‖ = 1 in units of δ, a = −(nC/δ) ĝ (s.t2), b = sN + nC ĝ (s.t1)
            const double ac = -(nC / delta);
            if (l < kN) {
                s.t2[l] = ac * s.r[l];
                s.t1[l] = 1.0 * s.d[l] + nC * s.r[l];
            }
            sync();
            const double binv = 1.0 / fold_rnorm(s.t1);
            if (l < kN) s.t1[l] = binv * s.t1[l];
            sync();
            const double ab = fold_dot(s.t2, s.t1);
            const double c = 1.0 - fold_norm2(s.t2);
            const double ab2 = ab * ab;
            const double t = c / (ab + __builtin_sqrt(ab2 + c));
            const double dt = delta * t;
            if (l < kN) s.t3[l] = delta * s.t2[l] + dt * s.t1[l];
            boundary = true;
        }
        sync();
        jmatvec(s, f, q.mu, s.t3, s.t1);
        if (l < kN) s.t1[l] = 1.0 * s.Fx[l] + 1.0 * s.t1[l];
        sync();
        const double pred = fnorm - fold_rnorm(s.t1);
        if (l < kN) s.t2[l] = s.x[l] + 1.0 * s.t3[l];
        sync();
        eval_F(s, f, s.t2, s.t4);
        const double fnp = fold_rnorm(s.t4);
        const double ared = fnorm - fnp;
        if (pred > 0.0 && ared >= 1e-4 * pred) {
            const double rho = ared / pred;
            if (rho < 0.1) {
                delta = nN < delta ? jmax(nN, lo) : dl_shrink(delta);
            } else if (rho > 0.75 && boundary) {
                delta = dl_grow(delta);
            }
            const double stepnorm = fold_rnorm(s.t3);
            if (l < kN) {
                s.x[l] = s.t2[l];
                s.Fx[l] = s.t4[l];
            }
            sync();
            st.radius = delta;
            Info r{kSuccess, fnp, stepnorm, 0, false};
            r.limited = !inside;
            return r;
        }
        if (delta == lo) {
            st.radius = delta;
            return Info{kLsFailed, fnorm, 0.0, 1, false};
        }
        delta = dl_shrink(delta);
    }
}

template <int GLOB> GS_DEV Info globalize(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st, bool exact) {
    if (GLOB == kBacktracking) return globalize_backtracking(s, f, q, st, exact);
    return globalize_static(s, f, q);
}

// The step's direction d = −(J \ F) into s.d from a register factor, then — when `refine` holds —
// refine_passes!! (refinement.jl:211-244) on it with the held Jreg: one loop whose pass 0 solves for F
// and whose later passes solve for the residual r = −F − J d, so that each factor type has one
// solve in the kernel. Returns refine_passes!!' `failed` (false without refinement).
// `mu` is the regularisation of the products: q.mu with a fresh factor (the regularised Jacobian it
// factors), 0 along a held factor (the JVP of F at x).
template <class T, bool REFINE>
GS_DEV bool direction_refined(MidShared& s, const Fpu& f, const DevParams& q, const T (&a)[kN], int pos, int fe,
                              bool refine, double mu) {
    const int l = lane_id();
    double fnorm = 0.0, rprev = 0.0;
    for (int it = 0;; ++it) {
        ldiv_rows(s, a, pos, fe, it == 0 ? s.Fx : s.r, it == 0 ? s.d : s.t2);
        if (it == 0) {
            if (l < kN) s.d[l] = -s.d[l];
            sync();
            if (!REFINE || !refine) return false;
            fnorm = fold_rnorm(s.Fx);
            if (fnorm == 0.0) return false;
            if (!isfin(fnorm)) return true;
            jmatvec(s, f, mu, s.d, s.t1);
            if (l < kN) s.r[l] = -s.Fx[l] + -s.t1[l];
            sync();
            rprev = fold_rnorm(s.r);
            if (!isfin(rprev)) return true;
            if (rprev <= q.ref_eta_eff * fnorm) return false;
            if (q.ref_maxiter < 1) return true;
        } else {
            if (l < kN) s.t3[l] = s.d[l] + s.t2[l];
            sync();
            jmatvec(s, f, mu, s.t3, s.t1);
            if (l < kN) s.r[l] = -s.Fx[l] + -s.t1[l];
            sync();
            const double rcur = fold_rnorm(s.r);
            if (!(rcur < rprev)) return true;
            if (l < kN) s.d[l] = s.t3[l];
            sync();
            rprev = rcur;
            if (rcur <= q.ref_eta_eff * fnorm) return false;
            if (it >= q.ref_maxiter) return true;
        }
    }
}

// The value factor_and_direction returns, and the code a step reports, when the rows are not tame
// and the kernel factors only tame rows: the system is left to the EXACT kernel.
constexpr int kRedoInfo = -1;
constexpr uint8_t kRedo = 0xFF;

// newton_factor's factorisation and the direction with T factors: Jreg at the state's point into
// register rows, the LU, then direction_refined. Returns the LU's info (no direction when > 0) or
// kRedoInfo; `failed` is direction_refined's. EXACT selects the LU with per-lane selects, exact for
// any rows; otherwise the zero-multiplier LU runs on tame rows and anything else is handed over.
// (One LU form per kernel: two inlined forms side by side cost the register allocator the factor's
// registers twice over.)
// The factor's rows, pivot map and scale go to (a, pos, fe): a held factor under reuse = k, or rows
// local to the step.
template <class T, bool REFINE, bool EXACT>
GS_DEV int factor_and_direction(MidShared& s, const Fpu& f, const DevParams& q, bool refine, bool& failed,
                                T (&a)[kN], int& pos, int& fe) {
    bool tame;
    fe = factor_rows<T>(s, f, q.mu, a, tame);
    failed = false;
    if (!EXACT && !tame) return kRedoInfo;
    const int info = lu_rows<T, !EXACT>(a, pos);
    failed = info == 0 && direction_refined<T, REFINE>(s, f, q, a, pos, fe, refine, q.mu);
    return info;
}
template <class T, bool REFINE, bool EXACT>
GS_DEV int factor_and_direction(MidShared& s, const Fpu& f, const DevParams& q, bool refine, bool& failed) {
    T a[kN];
    int pos, fe;
    return factor_and_direction<T, REFINE, EXACT>(s, f, q, refine, failed, a, pos, fe);
}

// newton_factor (newton.jl:1302-1311) up to the factorisation: J at x. Only the last
// factorisation's diagnostics reach the status, so the point and the ‖F‖ they read are kept and
// exit_diag forms them once (F and J are functions of x alone: the values are the same).
template <int GLOB> GS_DEV void factor_diag(MidShared& s, const Fpu& f, NewtonState& st) {
    st.jfnorm = state_fnorm<GLOB>(s, st);
    st.have_j = true;
    const int l = lane_id();
    if (l < kN) s.jx[l] = s.x[l];
    eval_K(s, f, s.x);
}

// grad_ratio_and_norm (newton.jl:1287-1298) of the last factorisation, at exit.
GS_DEV void exit_diag(MidShared& s, const Fpu& f, NewtonState& st) {
    if (!st.have_j) return;
    eval_K(s, f, s.jx);
    eval_F(s, f, s.jx, s.t1);
    grad_ratio(s, f, s.t1, st.jfnorm, st.gradratio, st.jnorm);
}

// newton_step!! plain (newton.jl:1679-1688).
// Under DogLeg (newton.jl:1725-1746, dogleg_direction): a failed factor, or a Newton step of zero or
// non-finite norm, leaves the Cauchy step.
template <bool REFINE, bool EXACT, int GLOB>
GS_DEV Info plain_step(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st) {
    factor_diag<GLOB>(s, f, st);
    bool failed;
    const int info = factor_and_direction<double, REFINE, EXACT>(s, f, q, true, failed);
    if (info == kRedoInfo) return Info{kRedo, 0.0, 0.0, 0, false};
    Info r;
    if constexpr (GLOB == kDogLeg) {
        bool newton = false;
        if (info == 0) {
            const double nd = fold_rnorm(s.d);
            newton = isfin(nd) && nd != 0.0;
        }
        r = dogleg_trial(s, f, q, st, newton);
    } else {
        r = info > 0 ? Info{kSingular, fold_rnorm(s.Fx), 0.0, 0, false}
                     : globalize<GLOB>(s, f, q, st, q.mu == 0.0 && !REFINE);
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

// mixed_step!! + mixed_newton!! with fresh factors (newton.jl:1604-1666, 1692-1720). The FP64
// factor serves both a state promoted before the step (with refinement) and the promotion after a
// failed FP32 step (promote_factor!!, newton.jl:1332-1339: no refinement), from one site.
template <bool REFINE, bool EXACT, int GLOB>
GS_DEV Info mixed_step(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st) {
    const double fnorm_old = state_fnorm<GLOB>(s, st);
    const bool was_promoted = st.promoted;
    bool ok = true;
    int32_t failures = 0;
    double lin_rel = 0.0;
    factor_diag<GLOB>(s, f, st);
    if (!was_promoted) {
        bool dir_failed;
        const int info32 = factor_and_direction<float, REFINE, EXACT>(s, f, q, true, dir_failed);
        if (info32 == kRedoInfo) return Info{kRedo, 0.0, 0.0, 0, false};
        const bool ref_failed = info32 > 0 || dir_failed;
        if (ref_failed) {
            failures += 1;
            st.promoted = true;
        } else {
            lin_rel = linear_residual_ratio(s, f, q, fnorm_old);
        }
    }
    if (st.promoted) {
        // The FP64 factor of a promoted state is left to the EXACT kernel, which solves such a
        // system again from its start: the fast kernel then holds only the FP32 factor's registers.
        if (!EXACT) return Info{kRedo, 0.0, 0.0, 0, false};
        bool failed;
        const int info64 = factor_and_direction<double, REFINE, EXACT>(s, f, q, was_promoted, failed);
        if (info64 == kRedoInfo) return Info{kRedo, 0.0, 0.0, 0, false};
        ok = info64 == 0;
    }
    const bool md_promoted = st.promoted;
    Info r;
    const bool promoted_step = md_promoted && !was_promoted;
    if constexpr (GLOB == kDogLeg) {
        // mixed_step!! with DogLeg (newton.jl:1757-1780): no factor leaves the Cauchy step
        bool newton = false;
        if (ok) {
            const double nd = fold_rnorm(s.d);
            newton = isfin(nd) && nd != 0.0;
        }
        r = dogleg_trial(s, f, q, st, newton);
    } else {
        if (!ok) {
            r = Info{kSingular, fold_rnorm(s.Fx), 0.0, failures, st.promoted};
            st.fnorm = r.fnorm;
            return r;
        }
        r = globalize<GLOB>(s, f, q, st, q.mu == 0.0 && md_promoted);
    }
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

// ---- reuse = k (newton.jl:1781-1874) --------------------------------------------------------------

// The factor the state holds, in the factor precision TF (its rows, pivot map and scale, the
// registers of a factorisation kept across steps), and the rule state FactorAge (newton.jl:621-633):
// its age, whether the last step stalled, whether the last step went along the held factor.
template <class TF> struct Held {
    TF a[kN];
    int pos, fe;
    int age;
    bool stalled, reused;
};

GS_DEV int fresh_age(const DevParams& q) { return q.reuse == 1 ? 0 : 1; }                    // newton.jl:629
GS_DEV int next_age(const DevParams& q, int age) { return age + 1 >= q.reuse ? 0 : age + 1; }  // newton.jl:632

// reuse_result (newton.jl:1814-1820): a SINGULAR step resets the age, a STALLED one refreshes the
// factor at the next step.
template <class TF> GS_DEV void set_rule(Held<TF>& h, uint8_t code, int age_next, bool along_held) {
    h.age = code == kSingular ? 0 : age_next;
    h.stalled = code == kStalled;
    h.reused = along_held;
}

// newton_step!! with reuse in the working precision (newton.jl:1781-1810): a step along the held
// factor refines with the JVP at x and takes the measured slope; a failed refinement (one step-rule
// failure) or a line search that fails without a step refactorises at x in the same step.
template <bool REFINE, bool EXACT, int GLOB>
GS_DEV Info plain_step_reuse(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st, Held<double>& h) {
    int32_t failures = 0;
    if (!(h.age == 0 || h.stalled)) {
        const int age_next = next_age(q, h.age);
        eval_K(s, f, s.x);
        if (direction_refined<double, REFINE>(s, f, q, h.a, h.pos, h.fe, true, 0.0)) {
            failures = 1;
        } else {
            // a failed search without a step leaves x, Fx and the merit as they were
            const Info r = globalize<GLOB>(s, f, q, st, false);
            if (!(r.code == kLsFailed && r.stepnorm == 0.0)) {
                st.fnorm = r.fnorm;
                set_rule(h, r.code, age_next, true);
                return r;
            }
        }
    }
    factor_diag<GLOB>(s, f, st);
    bool failed;
    const int info = factor_and_direction<double, REFINE, EXACT>(s, f, q, true, failed, h.a, h.pos, h.fe);
    if (info == kRedoInfo) return Info{kRedo, 0.0, 0.0, 0, false};
    Info r = info > 0 ? Info{kSingular, fold_rnorm(s.Fx), 0.0, 0, false}
                      : globalize<GLOB>(s, f, q, st, q.mu == 0.0 && !REFINE);
    r.failures += failures;
    st.fnorm = r.fnorm;
    set_rule(h, r.code, fresh_age(q), false);
    return r;
}

// mixed_step!! with reuse (newton.jl:1823-1874) over mixed_newton!! with store = true
// (newton.jl:1604-1666): a promoted state factorises in FP64 at every step (left to the EXACT
// kernel, as without reuse); otherwise the step runs along the held FP32 factor until its age or a
// stall refreshes it, and a held direction whose refinement fails is refactorised in FP32 at x (one
// failure) before any promotion. A line search along the held factor that fails without a step is
// retried once from a fresh factor.
template <bool REFINE, bool EXACT, int GLOB>
GS_DEV Info mixed_step_reuse(MidShared& s, const Fpu& f, const DevParams& q, NewtonState& st, Held<float>& h) {
    const double fnorm_old = state_fnorm<GLOB>(s, st);
    const bool state_promoted = st.promoted;
    bool fresh = st.promoted || h.age == 0 || h.stalled;
    const int age_fresh = fresh_age(q);
    int age_next = fresh ? age_fresh : next_age(q, h.age);
    while (true) {
        const bool was_fresh = fresh;
        bool ok = true, md_fresh = true, have = false;
        int32_t failures = 0;
        double lin_rel = 0.0;
        if (st.promoted) {
            if (!EXACT) return Info{kRedo, 0.0, 0.0, 0, false};
            factor_diag<GLOB>(s, f, st);
            bool failed;
            ok = factor_and_direction<double, REFINE, EXACT>(s, f, q, true, failed) == 0;
            have = true;
        } else if (!fresh) {
            eval_K(s, f, s.x);
            if (!direction_refined<float, REFINE>(s, f, q, h.a, h.pos, h.fe, true, 0.0)) {
                md_fresh = false;
                have = true;
            }
        }
        if (!have) {
            factor_diag<GLOB>(s, f, st);
            if (!fresh) failures += 1;
            bool dir_failed;
            const int info32 = factor_and_direction<float, REFINE, EXACT>(s, f, q, true, dir_failed, h.a, h.pos, h.fe);
            if (info32 == kRedoInfo) return Info{kRedo, 0.0, 0.0, 0, false};
            if (info32 > 0 || dir_failed) {
                // promote_factor!! (newton.jl:1332-1339): the held Jacobian in FP64, no refinement
                failures += 1;
                st.promoted = true;
                if (!EXACT) return Info{kRedo, 0.0, 0.0, 0, false};
                bool failed;
                ok = factor_and_direction<double, REFINE, EXACT>(s, f, q, false, failed) == 0;
            } else {
                lin_rel = linear_residual_ratio(s, f, q, fnorm_old);
            }
        }
        fresh = md_fresh;
        if (!ok) {
            set_rule(h, kSingular, 0, !md_fresh);
            Info r{kSingular, fold_rnorm(s.Fx), 0.0, failures, st.promoted};
            st.fnorm = r.fnorm;
            return r;
        }
        const bool promoted_step = st.promoted && !state_promoted;
        if (promoted_step) {
            age_next = 0;
        } else if (md_fresh && !was_fresh) {
            age_next = age_fresh;
        }
        // newton_stepkind (newton.jl:1244-1254)
        Info r = globalize<GLOB>(s, f, q, st, md_fresh && q.mu == 0.0 && st.promoted);
        if (!md_fresh && r.code == kLsFailed && r.stepnorm == 0.0) {
            fresh = true;
            age_next = age_fresh;
            continue;
        }
        r.failures += failures;
        if (promoted_step) {
            r.promoted = true;
            st.slow = 0;
        } else if (md_fresh && !st.promoted) {
            // read_mixed_stall (newton.jl:1437-1451), η_lin = max(√eps, 1e-4)
            if (r.fnorm > 0.5 * fnorm_old && lin_rel > 1e-4) {
                const int32_t slow = st.slow + 1;
                if (slow >= 2) {
                    st.promoted = true;
                    st.slow = 0;
                    r.failures += 1;
                    r.promoted = true;
                } else {
                    st.slow = slow;
                }
            } else {
                st.slow = 0;
            }
        }
        st.fnorm = r.fnorm;
        set_rule(h, r.code, age_next, !md_fresh);
        return r;
    }
}

// ---- the solve ----------------------------------------------------------------------------------

GS_DEV void store_status_mid(gs_status_f64* dst, uint8_t code, int32_t iterations, double fnorm, double stepnorm,
                             int32_t failures, bool promoted, double gradratio, double rate, bool contracted,
                             bool reused) {
    uint64_t* w = reinterpret_cast<uint64_t*>(dst);
    w[0] = uint64_t(code) | (uint64_t(uint32_t(iterations)) << 32);
    w[1] = __builtin_bit_cast(uint64_t, fnorm);
    w[2] = __builtin_bit_cast(uint64_t, stepnorm);
    w[3] = uint64_t(uint32_t(failures)) | (uint64_t(promoted ? 1 : 0) << 32);
    w[4] = uint64_t(uint32_t(-1));
    w[5] = __builtin_bit_cast(uint64_t, dnan());
    w[6] = __builtin_bit_cast(uint64_t, gradratio);
    w[7] = __builtin_bit_cast(uint64_t, rate);
    w[8] = uint64_t(contracted ? 1 : 0) | (uint64_t(reused ? 1 : 0) << 8);
}

// One wave per system. The kernel with EXACT = false factors with the zero-multiplier LU and hands a
// system whose rows are ever not tame, or (mixed) whose state is ever promoted to the FP64 factor, to
// the EXACT kernel, launched after it over the same grid:
// that system keeps its input untouched and redo[idx] = 1; every other system writes redo[idx] = 0,
// and the EXACT kernel solves only the systems marked 1, from the start.
template <bool MIXED, bool REFINE, bool EXACT, int GLOB, bool REUSE>
__global__ __launch_bounds__(64) void fpu_gl3_kernel(const MidLaunch L) {
    __shared__ MidShared s;
    const int l = lane_id();
    const int64_t i = blockIdx.x;
    const int64_t idx = L.order ? L.order[i] - 1 : i;
    if (EXACT && L.redo[idx] == 0) return;
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
    NewtonState st{fnorm0, dnan(), 0.0, false, 0, fnorm0, 0.0, false,
                   GLOB == kBacktracking ? fold_norm2(s.Fx) : dnan(), q.dl_radius};
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
    Held<typename std::conditional<MIXED, float, double>::type> h;
    h.pos = 0;
    h.fe = 0;
    h.age = 0;
    h.stalled = false;
    h.reused = false;
    while (!done) {
        const double rate_k = rate;
        Info info;
        if constexpr (REUSE) {
            if constexpr (MIXED) info = mixed_step_reuse<REFINE, EXACT, GLOB>(s, f, q, st, h);
            else info = plain_step_reuse<REFINE, EXACT, GLOB>(s, f, q, st, h);
        } else {
            info = MIXED ? mixed_step<REFINE, EXACT, GLOB>(s, f, q, st) : plain_step<REFINE, EXACT, GLOB>(s, f, q, st);
        }
        if (!EXACT && info.code == kRedo) {
            if (l == 0) L.redo[idx] = 1;
            return;
        }
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
        // observe_rate!! (newton.jl:1205-1210): a step of rate θ ≥ 1 along a held factor refreshes
        // the factor at the next step
        if (REUSE && h.reused && rate >= 1.0) h.stalled = true;
        if (code == kSingular || code == kNonfinite || code == kMaxIters) break;
        if (code == kLsFailed && info.stepnorm == 0.0) break;
        const double ftest = q.norm_max ? fold_rmax(s.Fx) : fnorm;
        if (it >= q.min_iterations && ftest <= jmax(q.f_abstol, q.f_reltol * ftest0)) {
            code = kSuccess;
            break;
        }
        // a stalled line search, or a step that a trust radius limited, passes only the residual test
        const bool steptested = code != kStalled && !info.limited;
        const double xnorm = steptested ? fold_rnorm(s.x) : dnan();
        const double xtol = jmax(q.x_abstol, q.x_reltol * xnorm);
        const bool past_min = it >= q.min_iterations;
        bool ctr = false;
        if (!info.shortened && q.rate_test && past_min) {
            const double th = jmax(rate, rate_k);
            ctr = stepnorm > 0.0 && th < 1.0 && th / (1.0 - th) * stepnorm <= q.kappa * xtol;
        }
        const bool step_conv = past_min && it > 0 && stepnorm > 0.0 && stepnorm <= xtol;
        const bool zero_step = code == kSuccess && info.stepnorm == 0.0 && past_min;
        const bool rstop = q.roundoff && past_min && rate >= 1.0 && rate_k < 1.0;
        if (ctr || (steptested && (step_conv || zero_step || rstop))) {
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
    if (L.S) exit_diag(s, f, st);
    if (l == 0) {
        if (!EXACT) L.redo[idx] = 0;
        if (L.S) store_status_mid(L.S + idx, code, it, fnorm, stepnorm, failures, promoted, st.gradratio, rate, contracted,
                                    REUSE && h.reused && it > 0);  // reused_factor (newton.jl:1188-1191)
        if (L.codes) L.codes[idx] = code;
    }
}

// The two launches of one form: the fast kernel, then the EXACT kernel over the systems it handed over.
template <bool MIXED, bool REFINE, int GLOB, bool REUSE> int launch_glob_variant(const MidLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL((fpu_gl3_kernel<MIXED, REFINE, false, GLOB, REUSE>), dim3(unsigned(L.nsys)), dim3(64), 0, st, L);
    hipLaunchKernelGGL((fpu_gl3_kernel<MIXED, REFINE, true, GLOB, REUSE>), dim3(unsigned(L.nsys)), dim3(64), 0, st, L);
    return GS_OK;
}

template <bool REUSE> int launch_forms(const MidLaunch& L, hipStream_t st) {
    if constexpr (!REUSE) {
        // DogLeg takes no reuse (newton.jl:66-70)
        if (L.glob == kDogLeg) {
            if (L.mixed)
                return L.refine ? launch_glob_variant<true, true, kDogLeg, false>(L, st)
                                : launch_glob_variant<true, false, kDogLeg, false>(L, st);
            return L.refine ? launch_glob_variant<false, true, kDogLeg, false>(L, st)
                            : launch_glob_variant<false, false, kDogLeg, false>(L, st);
        }
    }
    const bool bt = L.glob == kBacktracking;
    if (L.mixed) {
        if (L.refine)
            return bt ? launch_glob_variant<true, true, kBacktracking, REUSE>(L, st)
                      : launch_glob_variant<true, true, kStatic, REUSE>(L, st);
        return bt ? launch_glob_variant<true, false, kBacktracking, REUSE>(L, st)
                  : launch_glob_variant<true, false, kStatic, REUSE>(L, st);
    }
    if (L.refine)
        return bt ? launch_glob_variant<false, true, kBacktracking, REUSE>(L, st)
                  : launch_glob_variant<false, true, kStatic, REUSE>(L, st);
    return bt ? launch_glob_variant<false, false, kBacktracking, REUSE>(L, st)
              : launch_glob_variant<false, false, kStatic, REUSE>(L, st);
}

}  // namespace

}  // namespace gs
