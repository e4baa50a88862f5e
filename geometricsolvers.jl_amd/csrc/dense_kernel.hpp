// The kernels of the dense mid-size solve, included by the translation units that instantiate them
// (dense_n16.hip … dense_n64.hip, one size each, so that they compile in parallel); the launch
// dispatch and the shape setting are in dense_solver.hip.
//
// The §10.E problem of the reference's device benchmark (benchmark/device_problems.jl:133-175),
// F(x, b) = A x + sin.(x)/10 − b with A shared by the batch and b per system (GS_FAM_DENSE_SIN), under
// Newton with the Static or Backtracking globalization, with or without reuse = k, the static LU in
// FP64 or in FP32 with FP64 refinement and promotion: the reference's R1 loop solve!! (src/nonlinear/solver.jl:277-378) with
// newton_step!! (src/nonlinear/newton.jl:1302-1485, 1604-1720) per system, as the lane-per-system
// kernels run it for n ≤ 8.
//
// Layout. A system owns a segment of NP = 16, 32 or 64 lanes (four, two or one systems per wave;
// NP ≥ n, the launch's tunable choice, by default the smallest that holds the rows); lane i < n of
// the segment owns row i of the factor (registers: n doubles, or n floats for the FP32 factor) and
// entry i of every vector (LDS, one region per system). Up to n = 32 the lane also holds its row of
// the Jacobian; from n = 48 on that row is formed again from A in LDS where it is read. The scalar state of a solve is computed by every lane of its segment alike, so
// control flow is uniform within a segment; segments of one wave diverge under EXEC like lanes.
// Nothing crosses a segment: the pivot search is a segment reduction (DPP within the 16-lane rows, a
// swizzle across the two rows of a 32-lane segment and a permute across the halves of a 64-lane one), the pivot row and the triangular solves'
// finished entries pass through the system's LDS region, and the only workgroup barrier is the one
// after staging A in LDS. LDS traffic within a segment needs no barrier (a wave's LDS operations
// complete in order); a wave-scope fence keeps the compiler from moving a read across another
// lane's write.
//
// The factorisation is the reference's unblocked LU in the factor's precision (lu.jl:24-64): the
// first-largest pivot over the logical rows ≥ k (a NaN |a_kk| keeps row k), rows that never move
// (each lane carries the logical index `pos` of its row; a swap exchanges two indices), exact a/p
// multipliers and the rank-1 update without FMA, applied only to the rows below the pivot — so both
// factor precisions match the oracle bit for bit. Sums the reference folds left to right (norms,
// products, ‖J‖_F) are folded in that order by every lane over the LDS vector (broadcast reads).
//
// Diagnostics. The status's gradient ratio ‖Jᴴ F‖/(‖J‖_F ‖F‖) is that of the last factorisation
// (newton.jl:1287-1311, 1165-1231). It changes nothing else in a solve of a NonlinearProblem, so the
// kernel keeps the point of each factorisation (x, F and the ‖F‖ it read) and evaluates the ratio
// once, after the loop, from the last one.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "dense_solver.hpp"
#include "small_launch.hpp"

namespace gs {

int set_error(int code, const std::string& msg);

namespace {

// ---- the shape of a launch ------------------------------------------------------------------------

// NP lanes per system (16, 32 or 64, at least n), G = 64 / NP systems per wave, four waves per
// workgroup. The lanes per system are the launch's tunable choice (gs_dense_set_shape).
template <int N, int NP_> struct Shape {
    static_assert(N == 12 || N == 16 || N == 24 || N == 32 || N == 48 || N == 64,
                  "the dense solve takes n = 12, 16, 24, 32, 48 or 64");
    static_assert((NP_ == 16 || NP_ == 32 || NP_ == 64) && NP_ >= N, "a segment of 16, 32 or 64 lanes, one per row");
    static constexpr int NP = NP_;                // lanes per system
    static constexpr int G = 64 / NP;             // systems per wave
    static constexpr int kWaves = 4;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int SPB = G * kWaves;        // systems per workgroup
};
constexpr int kDenseThreads = 256;  // every shape's workgroup


// One system's LDS region in the working precision T: the vectors of the solve, the pivot-row and
// solve broadcast buffers.
template <class T, int N> struct alignas(16) SysVecs {
    T x[N], Fx[N], d[N], xt[N], Ft[N], r[N], t[N], u[N];
    T jx[N], jF[N];  // the point of the last factorisation and its residual (diagnostics)
    T bx[N];         // the best iterate (best = true)
    double rowd[N];  // the pivot row (FP64 factor), a finished solve entry, a column of J in FP64
    float rowf[N];   // the same for the FP32 factor and J in FP32
    double yd[N];
    float yf[N];
};

// ---- segment primitives ---------------------------------------------------------------------------

// The compiler's wave-scope fence: no LDS access moves across it (the hardware keeps a wave's LDS
// accesses in order, so no instruction is needed).
GS_DEV void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A value brought by DPP from another lane of the same row of 16.
template <int CTRL> GS_DEV int dpp(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }
template <int CTRL> GS_DEV double dpp(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(dpp<CTRL>(int(uint32_t(b))));
    const uint32_t hi = uint32_t(dpp<CTRL>(int(uint32_t(b >> 32))));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
// The value of lane l ^ 16 within a group of 32 (ds_swizzle, bit mode: and 0x1F, xor 0x10).
GS_DEV int swz16(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x401F); }
GS_DEV double swz16(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(swz16(int(uint32_t(b))));
    const uint32_t hi = uint32_t(swz16(int(uint32_t(b >> 32))));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// The value of lane l ^ 32 (ds_bpermute through the LDS crossbar; no LDS memory is touched).
GS_DEV int xor32(int v) { return __builtin_amdgcn_ds_bpermute(int(((threadIdx.x & 63u) ^ 32u) << 2), v); }
GS_DEV double xor32(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(xor32(int(uint32_t(b))));
    const uint32_t hi = uint32_t(xor32(int(uint32_t(b >> 32))));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// The maximum of a NaN-free value over the segment, in every lane of it: quad swaps, the half-row
// and row mirrors give each row of 16 its maximum; a 32-lane segment then takes its other row's,
// and a 64-lane one its other half's.
template <int NP> GS_DEV double seg_max(double v) {
    v = __builtin_fmax(v, dpp<0xB1>(v));   // quad_perm [1,0,3,2]
    v = __builtin_fmax(v, dpp<0x4E>(v));   // quad_perm [2,3,0,1]
    v = __builtin_fmax(v, dpp<0x141>(v));  // row_half_mirror
    v = __builtin_fmax(v, dpp<0x140>(v));  // row_mirror
    if constexpr (NP >= 32) v = __builtin_fmax(v, swz16(v));
    if constexpr (NP == 64) v = __builtin_fmax(v, xor32(v));
    return v;
}

// Whether any lane of the segment holds b.
template <int NP> GS_DEV bool seg_any(bool b) {
    const uint64_t m = __ballot(b);
    const int base = int(threadIdx.x & 63) & ~(NP - 1);
    const uint64_t seg = NP == 64 ? ~0ull : ((1ull << NP) - 1);
    return ((m >> base) & seg) != 0;
}

// The smallest 32-bit integer over the segment, in every lane of it (the DPP move folds into the min).
template <int CTRL> GS_DEV int dpp_min(int v) {
    return __builtin_amdgcn_update_dpp(int(0x7fffffff), v, CTRL, 0xF, 0xF, false);
}
template <int NP> GS_DEV int seg_min(int v) {
    v = __builtin_elementwise_min(v, dpp_min<0xB1>(v));
    v = __builtin_elementwise_min(v, dpp_min<0x4E>(v));
    v = __builtin_elementwise_min(v, dpp_min<0x141>(v));
    v = __builtin_elementwise_min(v, dpp_min<0x140>(v));
    if constexpr (NP >= 32) v = __builtin_elementwise_min(v, swz16(v));
    if constexpr (NP == 64) v = __builtin_elementwise_min(v, xor32(v));
    return v;
}

// The pivot choice of a segment: the largest key, and among the lanes holding it the smallest tag
// (the tag orders by logical row first); every lane receives the winner's tag. Two reductions: the
// maximum of the keys, then the minimum of the tags of the lanes that hold it.
template <int NP> GS_DEV int seg_argmax(double key, int tag) {
    const double kmax = seg_max<NP>(key);
    return seg_min<NP>(key == kmax ? tag : int(0x7fffffff));
}

// ---- folds over an LDS vector, evaluated by every lane of the segment -----------------------------

template <int N, class T> GS_DEV T vnorm2(const T* v) {
    T s = v[0] * v[0];
#pragma unroll 8
    for (int i = 1; i < N; ++i) s = s + v[i] * v[i];
    return s;
}
template <int N, class T> GS_DEV T vrmax(const T* v) {
    T m = tabs(v[0]);
#pragma unroll 8
    for (int i = 1; i < N; ++i) m = jmax(m, tabs(v[i]));
    return m;
}
// rnorm (reductions.jl:41-57)
template <int N, class T> GS_DEV T vrnorm(const T* v) {
    const T s = vnorm2<N>(v);
    if (__builtin_expect(isfin(s) && s >= T(N) * Lim<T>::min, 1)) return tsqrt(s);
    const T m = vrmax<N>(v);
    if (m == T(0) || !isfin(m)) return m;
    const int e = exponent_of(m);
    T t = ldexp2(v[0], -e);
    T acc = t * t;
#pragma unroll 8
    for (int i = 1; i < N; ++i) {
        t = ldexp2(v[i], -e);
        acc = acc + t * t;
    }
    return ldexp2(tsqrt(acc), e);
}
// rdot (reductions.jl:77-85)
template <int N, class T> GS_DEV T vdot(const T* a, const T* b) {
    T s = a[0] * b[0];
#pragma unroll 8
    for (int i = 1; i < N; ++i) s = s + a[i] * b[i];
    return s;
}

// The exponent of the largest |entry| over the segment's entries `a` (lanes with `live`), clamped to
// ±emax of T; 0 for a zero vector or one with a non-finite entry (max_exponent, factors.jl:833-843).
// The maximum runs in FP64 (exact for an FP32 entry).
template <int NP, class T> GS_DEV int seg_exponent(T a, bool live) {
    const double m0 = live ? double(tabs(a)) : 0.0;
    const bool nan_any = seg_any<NP>(live && isnan_(m0));
    const double m = seg_max<NP>(isnan_(m0) ? 0.0 : m0);
    if (nan_any || !isfin(m) || m == 0.0) return 0;
    const int e = exponent_of(m);
    constexpr int emax = Lim<T>::emax;
    return e < -emax ? -emax : (e > emax ? emax : e);
}

// ---- the family: A x + sin.(x)/10 − b ---------------------------------------------------------------

// TenEResidual (device_problems.jl:147-157): row i of F and of J = A + Diagonal(cos.(x)/10), the
// matrix-vector product as StaticArrays' muladd fold over the row.
template <class R, int NN> struct FamDenseSin {
    using T = R;
    static constexpr int N = NN;
    const T* A;  // LDS, column-major
    T b;         // this lane's entry of the system's b
    GS_DEV T F(const T* x, int i) const {
        T acc = A[i] * x[0];
#pragma unroll 8
        for (int j = 1; j < N; ++j) acc = tfma(A[i + j * N], x[j], acc);
        T sn, cs;
        fsincos(x[i], sn, cs);
        return (acc + sn / T(10)) - b;
    }
    // J_ij at x from the lane's own diagonal term dg = cos(x_i)/10 (Jdiag): column j in parallel.
    GS_DEV T Jdiag(const T* x, int i) const {
        T sn, cs;
        fsincos(x[i], sn, cs);
        return cs / T(10);
    }
    GS_DEV T Jentry(int i, int j, T dg) const {
        const T a = A[i + j * N];
        return j == i ? a + dg : a;
    }
    GS_DEV void Jrow(const T* x, int i, T (&row)[N]) const {
        T sn, cs;
        fsincos(x[i], sn, cs);
        const T c10 = cs / T(10);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const T a = A[i + j * N];
            row[j] = j == i ? a + c10 : a;
        }
    }
};

// The lane's row of Jreg = J(x) + μI. Up to n = 32 it is held in VGPRs (formed once per
// factorisation); from n = 48 on it would not fit beside the factor's row, so each entry is formed
// again from A in LDS and the lane's diagonal term wherever it is read, with the same operations
// (a + cos(x_i)/10, then + μ where μ ≠ 0), so the same bits.
template <int N> constexpr bool holds_jrow() { return N <= 32; }
template <class T, int N, bool HOLD = holds_jrow<N>()> struct JacRow {
    T r[N];
    GS_DEV T operator[](int j) const { return r[j]; }
};
template <class T, int N> struct JacRow<T, N, false> {
    const T* A;  // LDS, column-major
    int i;       // the row
    T dg, mu;    // cos(x_i)/10 and the regularization
    GS_DEV T operator[](int j) const {
        const T a = A[i + j * N];
        if (j != i) return a;
        const T d = a + dg;
        return mu != T(0) ? d + mu : d;
    }
};

// ---- the solve of one system --------------------------------------------------------------------

template <class Fam, int NP_, int GLOB, bool MIXED, bool REFINE, bool DIAG, bool REUSE = false> struct DenseSolve {
    static constexpr int N = Fam::N;
    static constexpr int NP = NP_;
    using T = typename Fam::T;  // the working precision
    static_assert(sizeof(T) == 8 || !MIXED, "a Float32 batch factorises in its working precision");
    using JR = JacRow<T, N>;
    using Step = StepT<T>;
    using TF = typename std::conditional<MIXED, float, T>::type;  // the factor's (and held factor's) type

    const Fam& fam;
    const DevParamsT<T>& q;
    SysVecs<T, N>& v;
    const int sl;      // the lane's row within the segment
    const bool live;   // sl < N

    // NewtonState (newton.jl:715-744): uniform within the segment
    T merit, sfnorm;
    bool promoted;
    int slow;
    T diag_fnorm;  // the ‖F‖ the last factorisation read; NaN before the first
    // reuse = k: the factor the state holds (this lane's row, its logical row and the prescale) and
    // the rule state FactorAge (newton.jl:621-633): its age, whether the last step stalled, whether
    // the last step went along the held factor
    TF ka[N];
    int kpos, kfe;
    int age;
    bool stalled, reused;
    // DogLeg: the trust radius (TrustRadius, newton.jl:637-640), 0 until the first step sets it
    T radius;
    static_assert(!(REUSE && GLOB == kDogLeg), "Newton with a DogLeg globalization takes no reuse");

    GS_DEV DenseSolve(const Fam& f, const DevParamsT<T>& p, SysVecs<T, N>& vv, int lane)
        : fam(f), q(p), v(vv), sl(lane), live(lane < N) {}

    // F at the LDS vector xin into out (every lane of the segment, then the fence).
    GS_DEV void eval_F(const T* xin, T* out) {
        if (live) out[sl] = fam.F(xin, sl);
        wsync();
    }

    GS_DEV void copy(const T* from, T* to) {
        if (live) to[sl] = from[sl];
        wsync();
    }

    // residual_norm (newton.jl:1274-1278)
    GS_DEV T residual_norm() const {
        if (GLOB == kBacktracking) {
            if (isfin(merit) && merit >= T(N) * Lim<T>::min) return tsqrt(merit);
            return vrnorm<N>(v.Fx);
        }
        return isnan_(sfnorm) ? vrnorm<N>(v.Fx) : sfnorm;
    }

    // ---- the Jacobian, its factor and the solves ----

    // newton_factor (newton.jl:1302-1311) up to the factorisation: Jreg's row at x, the point kept
    // for the diagnostics.
    GS_DEV void jacobian(JR& jr) {
        if (DIAG) {
            diag_fnorm = residual_norm();
            if (live) {
                v.jx[sl] = v.x[sl];
                v.jF[sl] = v.Fx[sl];
            }
        }
        if constexpr (holds_jrow<N>()) {
            fam.Jrow(v.x, live ? sl : 0, jr.r);
            if (q.mu != T(0)) {
#pragma unroll
                for (int j = 0; j < N; ++j) jr.r[j] = j == sl ? jr.r[j] + q.mu : jr.r[j];
            }
        } else {
            jr.A = fam.A;
            jr.i = live ? sl : 0;
            jr.dg = fam.Jdiag(v.x, jr.i);
            jr.mu = q.mu;
        }
    }

    // This lane's row of J at the LDS point xin, without the regularization (the JVP of F there).
    GS_DEV void jrow_at(const T* xin, JR& ja) {
        if constexpr (holds_jrow<N>()) {
            fam.Jrow(xin, live ? sl : 0, ja.r);
        } else {
            ja.A = fam.A;
            ja.i = live ? sl : 0;
            ja.dg = fam.Jdiag(xin, ja.i);
            ja.mu = T(0);
        }
    }

    template <class TB> GS_DEV TB* rowbuf() {
        if constexpr (sizeof(TB) == 8) return v.rowd;
        else return v.rowf;
    }
    template <class TB> GS_DEV TB* ybuf() {
        if constexpr (sizeof(TB) == 8) return v.yd;
        else return v.yf;
    }

    // factor_matrix (factors.jl:802-811) + factorize!! (lu.jl:24-64) on the segment's rows: a
    // receives this lane's row of Jreg/2^e in TB, LU-factored; pos its logical row; returns info (the
    // first zero pivot, 1-based, or 0); fe the prescale exponent (applied when ≠ 0 for a factor
    // narrower than T, only when < 0 in the working precision). The row maximum runs in FP64 (exact).
    template <class TB> GS_DEV int factor(const JR& jr, TB (&a)[N], int& pos, int& fe) {
        double mx = 0.0;
        bool nan_seen = false;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            mx = __builtin_fmax(mx, double(tabs(jr[j])));
            nan_seen |= isnan_(jr[j]);
        }
        const bool nan_any = seg_any<NP>(live && nan_seen);
        const double wmx = seg_max<NP>(live ? mx : 0.0);
        int e0 = 0;
        constexpr int emax = Lim<T>::emax;
        if (!nan_any && isfin(wmx) && wmx != 0.0) {
            const int ex = exponent_of(wmx);
            e0 = ex < -emax ? -emax : (ex > emax ? emax : ex);
        }
        constexpr bool same = sizeof(TB) == sizeof(T);
        fe = same ? (e0 < 0 ? e0 : 0) : e0;
#pragma unroll
        for (int j = 0; j < N; ++j) a[j] = TB(fe != 0 ? ldexp2(jr[j], -fe) : jr[j]);
        TB* rb = rowbuf<TB>();
        pos = live ? sl : 255;
        int info = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const bool cand = live && pos >= k;
            const double mag = double(tabs(a[k]));
            const double key = !cand ? -1.0 : (isnan_(mag) ? (pos == k ? double(__builtin_inf()) : -0.5) : mag);
            const int tag = seg_argmax<NP>(key, (pos << 8) | sl);
            const int pw = tag >> 8, lw = tag & 255;
            if (pos == k) pos = pw;
            else if (sl == lw) pos = k;
            if (sl == lw) {
#pragma unroll
                for (int j = k; j < N; ++j) rb[j] = a[j];
            }
            wsync();
            const TB ukk = rb[k];
            const bool zero = ukk == TB(0);
            if (zero && info == 0) info = k + 1;
            if (live && pos > k && !zero) {
                const TB lik = a[k] / ukk;
                a[k] = lik;
#pragma unroll
                for (int j = k + 1; j < N; ++j) a[j] = a[j] - lik * rb[j];
            }
            wsync();
        }
        return info;
    }

    // backsolve (lu.jl:67-92) on the segment's rows: y is this lane's entry of the right-hand side
    // (rows never moved); returns, in the lane of logical row p, entry p of the solution.
    template <class TB> GS_DEV TB backsolve(const TB (&a)[N], int pos, TB y) {
        TB* yb = ybuf<TB>();
#pragma unroll
        for (int k = 0; k < N - 1; ++k) {
            if (pos == k) yb[k] = y;
            wsync();
            const TB yk = yb[k];
            if (live && pos > k) y = y - a[k] * yk;
        }
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {
            if (pos == k) {
                y = y / a[k];
                yb[k] = y;
            }
            wsync();
            const TB xk = yb[k];
            if (live && pos < k) y = y - a[k] * xk;
        }
        return y;
    }

    // ldiv!! with factors in the working precision (factors.jl:931-936): out = J \ r.
    GS_DEV void ldiv_same(const T (&a)[N], int pos, int fe, const T* r, T* out) {
        const T ri = live ? r[sl] : T(0);
        const int m = seg_exponent<NP>(ri, live);
        const int lo = fe < 0 ? fe : 0;
        const int c = m < lo ? lo : (m > 0 ? 0 : m);
        const int alt = m - (Lim<T>::emax - headroom(N));
        const int k = c > alt ? c : alt;
        T y = k != 0 ? ldexp2(ri, -k) : ri;
        y = backsolve<T>(a, pos, y);
        const int e = k - fe;
        if (live) out[pos] = e != 0 ? ldexp2(y, e) : y;
        wsync();
    }

    template <class TB> GS_DEV void ldiv(const TB (&a)[N], int pos, int fe, const T* r, T* out) {
        if constexpr (sizeof(TB) == sizeof(T)) ldiv_same(a, pos, fe, r, out);
        else ldiv_mixed(a, pos, fe, r, out);
    }

    // ldiv!! with FP32 factors in an FP64 solve (factors.jl:943-949): r/s in FP32 with s = max|r|,
    // then a·frac·2^(ex−e) in FP64.
    GS_DEV void ldiv_mixed(const float (&a)[N], int pos, int fe, const double* r, double* out) {
        const double ri = live ? r[sl] : 0.0;
        const double ar = __builtin_fabs(ri);
        const bool nan_any = seg_any<NP>(live && isnan_(ar));
        const double sm = nan_any ? dnan() : seg_max<NP>(ar);
        if (sm == 0.0) {
            if (live) out[sl] = 0.0;
            wsync();
            return;
        }
        float y = float(ri / sm);
        y = backsolve<float>(a, pos, y);
        const int ex = isfin(sm) ? __builtin_amdgcn_frexp_exp(sm) : 0;
        const double frac = isfin(sm) ? __builtin_amdgcn_frexp_mant(sm) : sm;
        if (live) out[pos] = ldexp2(double(y) * frac, ex - fe);
        wsync();
    }

    // out = Jreg v (StaticArrays' muladd fold over the row).
    GS_DEV void matvec(const JR& jr, const T* vin, T* out) {
        T acc = jr[0] * vin[0];
#pragma unroll
        for (int j = 1; j < N; ++j) acc = tfma(jr[j], vin[j], acc);
        if (live) out[sl] = acc;
        wsync();
    }

    // The direction d = −(J \ F) from a factor in TF, then refine_passes!! (refinement.jl:211-244) with
    // the held Jreg; returns refine_passes!!' `failed` (false without refinement).
    template <class TB>
    GS_DEV bool direction(const JR& jr, const TB (&a)[N], int pos, int fe, bool refine) {
        ldiv(a, pos, fe, v.Fx, v.d);
        if (live) v.d[sl] = -v.d[sl];
        wsync();
        if (!REFINE || !refine) return false;
        const T fnorm = vrnorm<N>(v.Fx);
        if (fnorm == T(0)) return false;
        if (!isfin(fnorm)) return true;
        matvec(jr, v.d, v.t);
        if (live) v.r[sl] = -v.Fx[sl] + -v.t[sl];
        wsync();
        T rprev = vrnorm<N>(v.r);
        if (!isfin(rprev)) return true;
        if (rprev <= q.ref_eta_eff * fnorm) return false;
        for (int it = 1; it <= q.ref_maxiter; ++it) {
            ldiv(a, pos, fe, v.r, v.t);
            if (live) v.u[sl] = v.d[sl] + v.t[sl];
            wsync();
            matvec(jr, v.u, v.t);
            if (live) v.r[sl] = -v.Fx[sl] + -v.t[sl];
            wsync();
            const T rcur = vrnorm<N>(v.r);
            if (!(rcur < rprev)) return true;
            copy(v.u, v.d);
            rprev = rcur;
            if (rcur <= q.ref_eta_eff * fnorm) return false;
        }
        return true;
    }

    // ---- globalization ----

    // globalize!! Static (linefunction.jl:114-131)
    GS_DEV Step step_static() {
        const T dn = vrnorm<N>(v.d);
        const T stepnorm = tabs(q.alpha) * dn;
        if (!isfin(stepnorm)) return Step{kNonfinite, false, false, 0, vrnorm<N>(v.Fx), stepnorm, false};
        if (live) v.xt[sl] = v.x[sl] + q.alpha * v.d[sl];
        wsync();
        eval_F(v.xt, v.Ft);
        const T fn = vrnorm<N>(v.Ft);
        if (!isfin(fn)) return Step{kNonfinite, false, false, 0, vrnorm<N>(v.Fx), stepnorm, false};
        if (live) {
            v.x[sl] = v.xt[sl];
            v.Fx[sl] = v.Ft[sl];
        }
        wsync();
        return Step{kSuccess, false, false, 0, fn, stepnorm, false};
    }

    // globalize_merit!! for Backtracking (linefunction.jl:159-199 → linesearch.jl:154-171 →
    // backtracking.jl:354-390); `exact`: ExactStep (slope −2φ₀) or MeasuredSlope (φ′(0) = 2⟨F, J d⟩).
    GS_DEV Step step_backtracking(bool exact) {
        const T phi0 = merit;
        const T dn = vrnorm<N>(v.d);
        if (!isfin(dn)) return Step{kNonfinite, false, false, 0, vrnorm<N>(v.Fx), dn, false};
        T d0;
        if (exact) {
            d0 = -T(2) * phi0;
        } else {
            if (live) v.xt[sl] = v.x[sl] + T(0) * v.d[sl];
            wsync();
            eval_F(v.xt, v.Ft);
            JR ja;  // J at x + 0·d, without the regularization
            jrow_at(v.xt, ja);
            matvec(ja, v.d, v.t);
            d0 = T(2) * vdot<N>(v.Ft, v.t);
        }
        if (!(isfin(phi0) && isfin(d0) && d0 < T(0))) {
            const uint8_t c = !(isfin(phi0) && isfin(d0)) ? kNonfinite : (d0 > T(0) ? kLsFailed : kStalled);
            if (c == kLsFailed) return Step{kLsFailed, false, false, 1, vrnorm<N>(v.Fx), T(0), false};
            return Step{c, false, false, 0, vrnorm<N>(v.Fx), T(0), false};
        }
        const T tau = tfmax(T(4) * Lim<T>::eps * tabs(phi0), T(4) * Lim<T>::tiny);
        const T amin = jmax(tau / tabs(d0), Lim<T>::min);
        T alpha = tfmin(tfmax(T(1), amin), T(__builtin_inf()));
        T aa = alpha, pa = tnan<T>();
        T ap = tnan<T>(), pp = tnan<T>();
        int frozen = 0;
        uint8_t code = kLsFailed;
        for (int trips = 0; trips < q.bt_maxiter; ++trips) {
            aa = alpha;
            if (live) v.xt[sl] = v.x[sl] + alpha * v.d[sl];
            wsync();
            eval_F(v.xt, v.Ft);
            pa = vnorm2<N>(v.Ft);
            frozen = (pa == phi0 && alpha <= Lim<T>::sqrteps) ? frozen + 1 : 0;
            if (pa - phi0 <= jmin(T(0), q.c1 * alpha * d0 + tau)) {
                code = (isfin(pa) && pa - phi0 < -tau) ? kSuccess : (tabs(pa - phi0) <= tau ? kStalled : kLsFailed);
                break;
            } else if (frozen >= 2) {
                code = kStalled;
                break;
            } else if (alpha <= amin) {
                code = tabs(pa - phi0) <= tau ? kStalled : kLsFailed;
                break;
            }
            const T an = isfin(pa) ? backtrack_next<T>(phi0, d0, alpha, pa, ap, pp, q.p)
                                   : jmin(T(0.1) * alpha, tsqrt(alpha) * tsqrt(amin));
            ap = alpha;
            pp = pa;
            alpha = jmax(an, amin);
        }
        if (code == kSuccess || (code == kLsFailed && isfin(pa) && pa < phi0)) {
            if (live) {
                v.x[sl] = v.x[sl] + aa * v.d[sl];
                v.Fx[sl] = v.Ft[sl];
            }
            wsync();
            merit = pa;
            return Step{code, false, aa < T(1), code == kLsFailed ? 1 : 0, tsqrt(pa), aa * dn, false};
        }
        if (code == kLsFailed) return Step{kLsFailed, false, false, 1, vrnorm<N>(v.Fx), T(0), false};
        return Step{code == kStalled ? kStalled : kNonfinite, false, false, 0, vrnorm<N>(v.Fx), T(0), false};
    }

    GS_DEV Step globalize(bool exact) {
        if (GLOB == kStatic) return step_static();
        return step_backtracking(exact);
    }

    // DogLeg (dogleg.jl:116-206, newton.jl:1725-1780): the radius clamps of the reference.
    GS_DEV static T dl_clamp(T a) { return a > T(1e10) ? T(1e10) : (a < T(1e-6) ? T(1e-6) : a); }
    GS_DEV static T dl_shrink(T a) { return jmax(T(0.25) * a, T(1e-6)); }
    GS_DEV static T dl_grow(T a) { return jmin(T(4) * a, T(1e10)); }

    // Jreg's diagonal entry of this lane's row at x (a_kk + cos(x_k)/10, then + μ where μ ≠ 0): the
    // same operations as the row jacobian() forms, so the same bits.
    GS_DEV T jreg_diag() const {
        const int k = live ? sl : 0;
        const T dd = fam.A[k + k * N] + fam.Jdiag(v.x, k);
        return q.mu != T(0) ? dd + q.mu : dd;
    }

    // out = Jregᵀ (c v), this lane's entry k = sl: column k of Jreg is column k of A (contiguous in
    // LDS) with the lane's own diagonal entry dkk; a left fold of fused multiply-adds over j, each
    // term's c·v_j formed first (the lane-per-system dogleg_trial and the oracle's matvec_t).
    GS_DEV void jtmatvec(T dkk, T c, const T* vin, T* out) {
        const int k = live ? sl : 0;
        const T* col = fam.A + k * N;
        T acc = (k == 0 ? dkk : col[0]) * (c * vin[0]);
#pragma unroll 8
        for (int j = 1; j < N; ++j) acc = tfma(j == k ? dkk : col[j], c * vin[j], acc);
        if (live) out[sl] = acc;
        wsync();
    }

    // The trial loop of dogleg_step on the model matrix Jreg (the row jr, the lane's diagonal dkk)
    // with the Newton step v.d (read where `newton`): the Newton step inside the radius, else the
    // Cauchy point cut to the radius, else the point of norm δ on the segment between them; a trial
    // is accepted when pred > 0 and ared ≥ 1e-4 pred (unsquared norms), a rejection shrinks δ to
    // max(δ/4, 1e-6), and a rejection at 1e-6 fails. ĝ in v.r, the trial step in v.u, its point in
    // v.xt and residual in v.Ft, the products in v.t; x, Fx and the radius updated.
    GS_DEV Step dogleg_trial(const JR& jr, T dkk, bool newton) {
        const T lo = T(1e-6);
        const T fnorm = vrnorm<N>(v.Fx);
        if (fnorm == T(0)) return Step{kSuccess, false, false, 0, fnorm, T(0), false};
        const T nN = newton ? vrnorm<N>(v.d) : T(__builtin_inf());
        // the Cauchy direction from the unit residual: g̃ = Jregᵀ (F/‖F‖), ĝ = g̃/‖g̃‖
        const T finv = T(1) / fnorm;
        jtmatvec(dkk, finv, v.Fx, v.r);
        const T gnorm = vrnorm<N>(v.r);
        const T ginv = T(1) / gnorm;
        if (live) v.r[sl] = ginv * v.r[sl];
        wsync();
        matvec(jr, v.r, v.t);
        const T jnorm = vrnorm<N>(v.t);
        const bool cauchy = gnorm != T(0) && isfin(gnorm) && jnorm != T(0) && isfin(jnorm);
        const T nC = fnorm * ((gnorm / jnorm) / jnorm);
        T delta = radius;
        if (delta == T(0) && (newton || cauchy)) {
            const T n0 = newton ? nN : nC;
            delta = n0 < lo ? T(2) * lo : dl_clamp(n0);
        }
        while (true) {
            bool boundary;
            const bool inside = newton && nN <= delta;
            if (inside) {
                if (live) v.u[sl] = v.d[sl];
                boundary = nN >= delta;
            } else if (!cauchy) {
                radius = delta;
                return Step{kStalled, false, false, 0, fnorm, T(0), false};
            } else if (!newton || nC >= delta) {
                const T a = -jmin(nC, delta);
                if (live) v.u[sl] = a * v.r[sl];
                boundary = nC >= delta;
            } else {
                // the root t ≥ 0 of ‖a + t b̂‖ = 1 in units of δ, a = −(nC/δ) ĝ (v.u), b = sN + nC ĝ (v.t)
                const T ac = -(nC / delta);
                if (live) {
                    v.u[sl] = ac * v.r[sl];
                    v.t[sl] = T(1) * v.d[sl] + nC * v.r[sl];
                }
                wsync();
                const T binv = T(1) / vrnorm<N>(v.t);
                if (live) v.t[sl] = binv * v.t[sl];
                wsync();
                const T ab = vdot<N>(v.u, v.t);
                const T c = T(1) - vnorm2<N>(v.u);
                const T ab2 = ab * ab;
                const T t = c / (ab + tsqrt(ab2 + c));
                const T dt = delta * t;
                if (live) v.u[sl] = delta * v.u[sl] + dt * v.t[sl];
                boundary = true;
            }
            wsync();
            matvec(jr, v.u, v.t);
            if (live) {
                v.t[sl] = T(1) * v.Fx[sl] + T(1) * v.t[sl];
                v.xt[sl] = v.x[sl] + T(1) * v.u[sl];
            }
            wsync();
            const T pred = fnorm - vrnorm<N>(v.t);
            eval_F(v.xt, v.Ft);
            const T fnp = vrnorm<N>(v.Ft);
            const T ared = fnorm - fnp;
            if (pred > T(0) && ared >= T(1e-4) * pred) {
                const T rho = ared / pred;
                if (rho < T(0.1)) {
                    delta = nN < delta ? jmax(nN, lo) : dl_shrink(delta);
                } else if (rho > T(0.75) && boundary) {
                    delta = dl_grow(delta);
                }
                const T stepnorm = vrnorm<N>(v.u);
                if (live) {
                    v.x[sl] = v.xt[sl];
                    v.Fx[sl] = v.Ft[sl];
                }
                wsync();
                radius = delta;
                return Step{kSuccess, false, false, 0, fnp, stepnorm, !inside};
            }
            if (delta == lo) {
                radius = delta;
                return Step{kLsFailed, false, false, 1, fnorm, T(0), false};
            }
            delta = dl_shrink(delta);
        }
    }

    // A Newton step of finite, nonzero norm in v.d (dogleg_direction, newton.jl:1741-1746).
    GS_DEV bool newton_usable() const {
        const T nd = vrnorm<N>(v.d);
        return isfin(nd) && nd != T(0);
    }

    // ---- steps ----

    // newton_step!! in the working precision (newton.jl:1679-1688).
    GS_DEV Step plain_step() {
        JR jr;
        jacobian(jr);
        T a[N];
        int pos, fe;
        const int info = factor<T>(jr, a, pos, fe);
        Step s;
        if constexpr (GLOB == kDogLeg) {
            // newton_step!! with DogLeg (newton.jl:1725-1737): a failed factor leaves the Cauchy step
            bool newton = false;
            if (info <= 0) {
                direction<T>(jr, a, pos, fe, true);
                newton = newton_usable();
            }
            s = dogleg_trial(jr, jreg_diag(), newton);
        } else if (info > 0) {
            s = Step{kSingular, false, false, 0, vrnorm<N>(v.Fx), T(0), false};
        } else {
            direction<T>(jr, a, pos, fe, true);
            s = globalize(q.mu == T(0) && !REFINE);
        }
        sfnorm = s.fnorm;
        return s;
    }

    // mixed_step!! / mixed_newton!! with fresh factors (newton.jl:1604-1666, 1692-1720): the FP32
    // factor and its refinement, or the FP64 one of a promoted state; a failed FP32 factor or
    // refinement promotes and refactorises the held Jacobian in FP64 (promote_factor!!, newton.jl:
    // 1332-1339).
    GS_DEV Step mixed_step() {
        const double fnorm_old = residual_norm();
        const bool was_promoted = promoted;
        JR jr;
        jacobian(jr);
        int failures = 0;
        double lin_rel = 0.0;
        bool ok = true;
        if (!promoted) {
            float a[N];
            int pos, fe;
            const int info = factor<float>(jr, a, pos, fe);
            const bool ref_failed = info > 0 || direction<float>(jr, a, pos, fe, true);
            if (ref_failed) {
                failures += 1;
                promoted = true;
            } else if (fnorm_old != 0.0 && isfin(fnorm_old)) {
                // linear_residual_ratio (newton.jl:1422-1430)
                matvec(jr, v.d, v.t);
                if (live) v.u[sl] = 1.0 * v.Fx[sl] + 1.0 * v.t[sl];
                wsync();
                lin_rel = vrnorm<N>(v.u) / fnorm_old;
            }
        }
        if (promoted) {
            double a[N];
            int pos, fe;
            const int info = factor<double>(jr, a, pos, fe);
            if (info > 0) ok = false;
            else direction<double>(jr, a, pos, fe, was_promoted);
        }
        const bool promoted_step = promoted && !was_promoted;
        Step s;
        if constexpr (GLOB == kDogLeg) {
            // mixed_step!! with DogLeg (newton.jl:1757-1780): no factor leaves the Cauchy step
            s = dogleg_trial(jr, jreg_diag(), ok && newton_usable());
        } else {
            if (!ok) {
                s = Step{kSingular, promoted, false, failures, vrnorm<N>(v.Fx), 0.0, false};
                sfnorm = s.fnorm;
                return s;
            }
            s = globalize(q.mu == 0.0 && promoted);
        }
        s.failures += failures;
        if (promoted_step) {
            s.promoted = true;
            slow = 0;
        } else if (!was_promoted) {
            // read_mixed_stall (newton.jl:1437-1451), η_lin = max(√eps, 1e-4)
            if (s.fnorm > 0.5 * fnorm_old && lin_rel > 1e-4) {
                const int sw = slow + 1;
                if (sw >= 2) {
                    promoted = true;
                    slow = 0;
                    s.failures += 1;
                    s.promoted = true;
                } else {
                    slow = sw;
                }
            } else {
                slow = 0;
            }
        }
        sfnorm = s.fnorm;
        return s;
    }

    // ---- reuse = k (newton.jl:1781-1874), as the lane-per-system reuse steps ----

    GS_DEV int fresh_age() const { return q.reuse == 1 ? 0 : 1; }                // newton.jl:629
    GS_DEV int next_age() const { return age + 1 >= q.reuse ? 0 : age + 1; }      // newton.jl:632

    // reuse_result (newton.jl:1814-1820): a SINGULAR step resets the age, a STALLED one refreshes
    // the factor at the next step.
    GS_DEV void set_rule(uint8_t code, int age_next, bool along_held) {
        age = code == kSingular ? 0 : age_next;
        stalled = code == kStalled;
        reused = along_held;
    }

    // newton_step!! with reuse in the working precision (newton.jl:1781-1810): a step along the held
    // factor refines with the JVP at x and takes the measured slope; a failed refinement (one
    // step-rule failure) or a line search that fails without a step refactorises at x in the step.
    GS_DEV Step plain_step_reuse() {
        int failures = 0;
        if (!(age == 0 || stalled)) {
            const int age_next = next_age();
            JR ja;
            jrow_at(v.x, ja);
            if (direction<TF>(ja, ka, kpos, kfe, true)) {
                failures = 1;
            } else {
                // a failed search without a step leaves x, Fx and the merit as they were
                Step s = globalize(false);
                if (!(s.code == kLsFailed && s.stepnorm == T(0))) {
                    sfnorm = s.fnorm;
                    set_rule(s.code, age_next, true);
                    return s;
                }
            }
        }
        JR jr;
        jacobian(jr);
        const int info = factor<TF>(jr, ka, kpos, kfe);
        Step s;
        if (info > 0) {
            s = Step{kSingular, false, false, 0, vrnorm<N>(v.Fx), T(0), false};
        } else {
            direction<TF>(jr, ka, kpos, kfe, true);
            s = globalize(q.mu == T(0) && !REFINE);
        }
        s.failures += failures;
        sfnorm = s.fnorm;
        set_rule(s.code, fresh_age(), false);
        return s;
    }

    // mixed_step!! with reuse (newton.jl:1823-1874) over mixed_newton!! with store = true
    // (newton.jl:1604-1666): a promoted state factorises in FP64 at every step; otherwise the step
    // runs along the held FP32 factor until its age or a stall refreshes it, and a held direction
    // whose refinement fails is refactorised in FP32 at x (one failure) before any promotion. A line
    // search along the held factor that fails without a step is retried once from a fresh factor.
    GS_DEV Step mixed_step_reuse() {
        const double fnorm_old = residual_norm();
        const bool state_promoted = promoted;
        bool fresh = promoted || age == 0 || stalled;
        const int age_fresh = fresh_age();
        int age_next = fresh ? age_fresh : next_age();
        while (true) {
            const bool was_fresh = fresh;
            bool ok = true, md_fresh = true, have = false;
            int failures = 0;
            double lin_rel = 0.0;
            JR jr;
            if (promoted) {
                jacobian(jr);
                double a[N];
                int pos, fe;
                if (factor<double>(jr, a, pos, fe) > 0) ok = false;
                else direction<double>(jr, a, pos, fe, true);
                have = true;
            } else if (!fresh) {
                JR ja;
                jrow_at(v.x, ja);
                if (!direction<TF>(ja, ka, kpos, kfe, true)) {
                    md_fresh = false;
                    have = true;
                }
            }
            if (!have) {
                jacobian(jr);
                if (!fresh) failures += 1;
                const int info = factor<TF>(jr, ka, kpos, kfe);
                if (info > 0 || direction<TF>(jr, ka, kpos, kfe, true)) {
                    // promote_factor!! (newton.jl:1332-1339): the held Jacobian in FP64, no refinement
                    failures += 1;
                    promoted = true;
                    double a[N];
                    int pos, fe;
                    if (factor<double>(jr, a, pos, fe) > 0) ok = false;
                    else direction<double>(jr, a, pos, fe, false);
                } else if (fnorm_old != 0.0 && isfin(fnorm_old)) {
                    // linear_residual_ratio (newton.jl:1422-1430)
                    matvec(jr, v.d, v.t);
                    if (live) v.u[sl] = 1.0 * v.Fx[sl] + 1.0 * v.t[sl];
                    wsync();
                    lin_rel = vrnorm<N>(v.u) / fnorm_old;
                }
            }
            fresh = md_fresh;
            if (!ok) {
                set_rule(kSingular, 0, !md_fresh);
                return Step{kSingular, promoted, false, failures, vrnorm<N>(v.Fx), 0.0, false};
            }
            const bool promoted_step = promoted && !state_promoted;
            if (promoted_step) {
                age_next = 0;
            } else if (md_fresh && !was_fresh) {
                age_next = age_fresh;
            }
            // newton_stepkind (newton.jl:1244-1254)
            Step s = globalize(md_fresh && q.mu == 0.0 && promoted);
            if (!md_fresh && s.code == kLsFailed && s.stepnorm == 0.0) {
                fresh = true;
                age_next = age_fresh;
                continue;
            }
            s.failures += failures;
            if (promoted_step) {
                s.promoted = true;
                slow = 0;
            } else if (md_fresh && !promoted) {
                // read_mixed_stall (newton.jl:1437-1451), η_lin = max(√eps, 1e-4)
                if (s.fnorm > 0.5 * fnorm_old && lin_rel > 1e-4) {
                    const int sw = slow + 1;
                    if (sw >= 2) {
                        promoted = true;
                        slow = 0;
                        s.failures += 1;
                        s.promoted = true;
                    } else {
                        slow = sw;
                    }
                } else {
                    slow = 0;
                }
            }
            sfnorm = s.fnorm;
            set_rule(s.code, age_next, !md_fresh);
            return s;
        }
    }

    GS_DEV Step step() {
        if constexpr (REUSE) {
            if constexpr (MIXED) return mixed_step_reuse();
            else return plain_step_reuse();
        } else {
            if constexpr (MIXED) return mixed_step();
            else return plain_step();
        }
    }

    // ---- the diagnostics of the last factorisation (grad_ratio_and_norm, newton.jl:1287-1298) ----

    // One pass over the columns of J at the saved point, column j formed in parallel (lane i its
    // entry J_ij) in v.rowd: the left fold of the squares in column-major order (rnorm's sum; with
    // e ≠ 0 of (J_ij/2^e)²) and the NaN-propagating largest |J_ij|, and, with `prod`, g = Jᴴ (sc F) into
    // v.r (entry j the muladd fold over column j).
    GS_DEV void diag_pass(T dg, int e, bool prod, T sc, T& sumsq, T& mx) {
        T s = T(0), m = T(0);
        T* col = rowbuf<T>();
#pragma unroll 1
        for (int j = 0; j < N; ++j) {
            if (live) col[sl] = fam.Jentry(sl, j, dg);
            wsync();
            T acc = T(0);
#pragma unroll 8
            for (int i = 0; i < N; ++i) {
                const T w = col[i];
                const T ws = e != 0 ? ldexp2(w, -e) : w;
                s = (j == 0 && i == 0) ? ws * ws : s + ws * ws;
                m = jmax(m, tabs(w));
                const T fi = sc != T(1) ? v.jF[i] * sc : v.jF[i];
                acc = i == 0 ? w * fi : tfma(w, fi, acc);
            }
            if (prod && sl == j) v.r[j] = acc;
            wsync();
        }
        sumsq = s;
        mx = m;
    }

    GS_DEV T grad_ratio() {
        if (isnan_(diag_fnorm)) return tnan<T>();
        const T dg = live ? fam.Jdiag(v.jx, sl) : T(0);
        const T fnorm = diag_fnorm;
        T s, m;
        diag_pass(dg, 0, true, T(1), s, m);
        T jnorm;
        if (__builtin_expect(isfin(s) && s >= T(N * N) * Lim<T>::min, 1)) {
            jnorm = tsqrt(s);
        } else if (m == T(0) || !isfin(m)) {
            jnorm = m;
        } else {
            const int e = exponent_of(m);
            T s2, m2;
            diag_pass(dg, e, false, T(1), s2, m2);
            jnorm = ldexp2(tsqrt(s2), e);
        }
        if (jnorm == T(0) || fnorm == T(0)) return T(0);
        const T gnorm = vrnorm<N>(v.r);
        if (isfin(gnorm) || !(isfin(jnorm) && isfin(fnorm))) return gnorm / jnorm / fnorm;
        const int e = exponent_of(fnorm);
        T s3, m3;
        diag_pass(dg, 0, true, ldexp2(T(1), -e), s3, m3);
        return vrnorm<N>(v.r) / jnorm / ldexp2(fnorm, -e);
    }

    // ---- the loop (solver.jl:277-378), as SmallSolve::solve ----

    struct Result {
        uint8_t code;
        bool contracted, promoted, reused;
        int32_t iterations, failures;
        T fnorm, stepnorm, rate, gradratio;
    };

    GS_DEV Result solve() {
        eval_F(v.x, v.Fx);
        const T fnorm0 = vrnorm<N>(v.Fx);
        merit = GLOB == kBacktracking ? vnorm2<N>(v.Fx) : tnan<T>();
        sfnorm = fnorm0;
        promoted = false;
        slow = 0;
        diag_fnorm = tnan<T>();
        kpos = 0;
        kfe = 0;
        age = 0;
        stalled = false;
        reused = false;
        radius = q.dl_radius;  // reset_rule (newton.jl:677)
        T best_fn = fnorm0;
        if (q.best) copy(v.x, v.bx);
        Result R{kSuccess, false, false, false, 0, 0, fnorm0, T(0), tnan<T>(), tnan<T>()};
        if (!isfin(fnorm0)) {
            R.code = kNonfinite;
        } else {
            const T ftest0 = q.norm_max ? vrmax<N>(v.Fx) : fnorm0;
            const T ftol = jmax(q.f_abstol, q.f_reltol * ftest0);
            if (0 >= q.min_iterations && ftest0 <= ftol) {
                R.code = kSuccess;
            } else if (q.max_iterations == 0) {
                R.code = kMaxIters;
            } else {
                int stalls = 0;
                while (true) {
                    const T rate_k = R.rate;
                    Step s = step();
                    if (q.best && s.fnorm < best_fn) {
                        best_fn = s.fnorm;
                        copy(v.x, v.bx);
                    }
                    R.rate = R.stepnorm > T(0) ? s.stepnorm / R.stepnorm : tnan<T>();
                    R.code = s.code;
                    R.iterations += 1;
                    R.fnorm = s.fnorm;
                    R.stepnorm = s.stepnorm;
                    R.failures += s.failures;
                    R.promoted |= s.promoted;
                    // observe_rate!! (newton.jl:1205-1210): a step of rate θ ≥ 1 along a held factor
                    // refreshes the factor at the next step
                    if (REUSE && reused && R.rate >= T(1)) stalled = true;
                    if (s.code == kSingular || s.code == kNonfinite || s.code == kMaxIters) break;
                    if (s.code == kLsFailed && s.stepnorm == T(0)) break;
                    const T ftest = q.norm_max ? vrmax<N>(v.Fx) : R.fnorm;
                    if (R.iterations >= q.min_iterations && ftest <= ftol) {
                        R.code = kSuccess;
                        break;
                    }
                    // a stalled line search, or a step that a trust radius limited, passes only the
                    // residual test
                    const bool steptested = s.code != kStalled && !s.limited;
                    const T xnorm = steptested ? vrnorm<N>(v.x) : tnan<T>();
                    const T xtol = jmax(q.x_abstol, q.x_reltol * xnorm);
                    const bool past_min = R.iterations >= q.min_iterations;
                    bool contracted = false;
                    if (!s.shortened && q.rate_test && past_min) {
                        const T th = jmax(R.rate, rate_k);
                        contracted = R.stepnorm > T(0) && th < T(1) && th / (T(1) - th) * R.stepnorm <= q.kappa * xtol;
                    }
                    const bool step_conv = past_min && R.stepnorm > T(0) && R.stepnorm <= xtol;
                    const bool zero_step = s.code == kSuccess && s.stepnorm == T(0) && past_min;
                    const bool roundoff = q.roundoff && past_min && R.rate >= T(1) && rate_k < T(1);
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
        // restore_best (newton.jl:792-799)
        if (q.best && R.code != kSuccess && isfin(best_fn) && best_fn < R.fnorm) {
            copy(v.bx, v.x);
            R.fnorm = best_fn;
        }
        if (DIAG) R.gradratio = grad_ratio();
        R.reused = REUSE && reused && R.iterations > 0;  // reused_factor (newton.jl:1188-1191)
        return R;
    }
};

// The status record of one system in the layout of SolverStatus{T}.
template <class T>
GS_DEV void store_status_t(typename StatusWords<T>::Status* dst, uint8_t code, int32_t iterations, T fnorm, T stepnorm,
                           int32_t failures, bool promoted, T gradratio, T rate, bool contracted, bool reused) {
    typename StatusWords<T>::W w[StatusWords<T>::kCount];
    pack_status(w, code, iterations, fnorm, stepnorm, failures, promoted, gradratio, rate, contracted, reused);
    store_status_lane(dst, w);
}

// Waves per SIMD the register allocator targets: 2 for the FP64 kernels at n = 64 beyond the plain
// Static form (Backtracking, DogLeg, the mixed and refined forms), which need a few registers more
// than 256 and would otherwise run at one wave per SIMD with the excess in AGPRs. Held to two waves
// they spill 200-850 bytes per lane to scratch and ran faster for 2^18 systems: Backtracking 98.3 ->
// 64.8 ms, DogLeg 112.4 -> 77.5 ms, the mixed form with refinement 103.0 -> 87.8 ms. The plain
// Static forms (with or without reuse) fit in 248 registers and take no target.
template <class T, int N, int GLOB, bool MIXED, bool REFINE> constexpr int dense_waves_per_eu() {
    return (sizeof(T) == 8 && N == 64 && (GLOB != kStatic || MIXED || REFINE)) ? 2 : 1;
}

template <class T, int N, int NP, int GLOB, bool MIXED, bool REFINE, bool DIAG, bool REUSE>
__global__ __launch_bounds__(kDenseThreads)
__attribute__((amdgpu_waves_per_eu(dense_waves_per_eu<T, N, GLOB, MIXED, REFINE>())))
void dense_kernel(const DenseLaunchT<T> L) {
    using Sh = Shape<N, NP>;
    __shared__ T sA[N * N];
    __shared__ SysVecs<T, N> sv[Sh::SPB];
    for (int t = int(threadIdx.x); t < N * N; t += Sh::kThreads) sA[t] = L.A[t];
    __syncthreads();
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    const int g = lane / Sh::NP, sl = lane & (Sh::NP - 1);
    const int sw = wave * Sh::G + g;
    const int64_t i = int64_t(blockIdx.x) * Sh::SPB + sw;
    if (i >= L.nsys) return;  // the whole segment leaves together
    const int64_t idx = L.order ? L.order[i] - 1 : i;
    SysVecs<T, N>& v = sv[sw];
    FamDenseSin<T, N> fam;
    fam.A = sA;
    const bool live = sl < N;
    fam.b = live ? reinterpret_cast<const T*>(L.P + (L.p_stride ? idx * L.p_stride : 0))[sl] : T(0);
    T* xp = L.X + idx * N;
    if (live) v.x[sl] = xp[sl];
    wsync();
    DenseSolve<FamDenseSin<T, N>, NP, GLOB, MIXED, REFINE, DIAG, REUSE> solver(fam, L.q, v, sl);
    const auto R = solver.solve();
    if (live) xp[sl] = v.x[sl];
    if (sl == 0) {
        if (DIAG) {
            store_status_t<T>(L.S + idx, R.code, R.iterations, R.fnorm, R.stepnorm, R.failures, R.promoted,
                              R.gradratio, R.rate, R.contracted, R.reused);
        } else {
            L.codes[idx] = R.code;
        }
    }
}

template <class T, int N, int NP, int GLOB, bool MIXED, bool REFINE, bool REUSE>
int launch_diag(bool diag, const DenseLaunchT<T>& L, hipStream_t st) {
    using Sh = Shape<N, NP>;
    const int64_t blocks = (L.nsys + Sh::SPB - 1) / Sh::SPB;
    if (blocks > 0x7fffffffLL) return set_error(GS_EINVAL, "batch too large for one launch");
    if (diag)
        hipLaunchKernelGGL((dense_kernel<T, N, NP, GLOB, MIXED, REFINE, true, REUSE>), dim3(unsigned(blocks)),
                           dim3(Sh::kThreads), 0, st, L);
    else
        hipLaunchKernelGGL((dense_kernel<T, N, NP, GLOB, MIXED, REFINE, false, REUSE>), dim3(unsigned(blocks)),
                           dim3(Sh::kThreads), 0, st, L);
    return GS_OK;
}

// The Float32 launches of one size and segment width (no mixed form), with and without reuse = k
// (dense_f16.hip … dense_f64.hip).
template <int N, int NP>
int launch_np_f32(int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L, hipStream_t st) {
    if (glob == kDogLeg) {
        if (reuse) return set_error(GS_EINVAL, "Newton with a DogLeg globalization takes no reuse");
        return refine ? launch_diag<float, N, NP, kDogLeg, false, true, false>(diag, L, st)
                      : launch_diag<float, N, NP, kDogLeg, false, false, false>(diag, L, st);
    }
    if (reuse) {
        if (glob == kBacktracking)
            return refine ? launch_diag<float, N, NP, kBacktracking, false, true, true>(diag, L, st)
                          : launch_diag<float, N, NP, kBacktracking, false, false, true>(diag, L, st);
        return refine ? launch_diag<float, N, NP, kStatic, false, true, true>(diag, L, st)
                      : launch_diag<float, N, NP, kStatic, false, false, true>(diag, L, st);
    }
    if (glob == kBacktracking)
        return refine ? launch_diag<float, N, NP, kBacktracking, false, true, false>(diag, L, st)
                      : launch_diag<float, N, NP, kBacktracking, false, false, false>(diag, L, st);
    return refine ? launch_diag<float, N, NP, kStatic, false, true, false>(diag, L, st)
                  : launch_diag<float, N, NP, kStatic, false, false, false>(diag, L, st);
}

// The launches of one size and segment width; REUSE: the reuse = k forms (their own translation
// units, dense_r16.hip … dense_r64.hip).
template <int N, int NP, bool REUSE = false>
int launch_np(int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st) {
    if (glob != kStatic && glob != kBacktracking)
        return set_error(GS_EINVAL, "this dense launch takes the Static or Backtracking globalization");
    if (glob == kBacktracking) {
        if (mixed) return refine ? launch_diag<double, N, NP, kBacktracking, true, true, REUSE>(diag, L, st)
                                 : launch_diag<double, N, NP, kBacktracking, true, false, REUSE>(diag, L, st);
        return refine ? launch_diag<double, N, NP, kBacktracking, false, true, REUSE>(diag, L, st)
                      : launch_diag<double, N, NP, kBacktracking, false, false, REUSE>(diag, L, st);
    }
    if (mixed) return refine ? launch_diag<double, N, NP, kStatic, true, true, REUSE>(diag, L, st)
                             : launch_diag<double, N, NP, kStatic, true, false, REUSE>(diag, L, st);
    return refine ? launch_diag<double, N, NP, kStatic, false, true, REUSE>(diag, L, st)
                  : launch_diag<double, N, NP, kStatic, false, false, REUSE>(diag, L, st);
}

// The DogLeg launches of one size and segment width (dense_d12.hip … dense_d64.hip).
template <int N, int NP>
int launch_np_dogleg(bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st) {
    if (mixed) return refine ? launch_diag<double, N, NP, kDogLeg, true, true, false>(diag, L, st)
                             : launch_diag<double, N, NP, kDogLeg, true, false, false>(diag, L, st);
    return refine ? launch_diag<double, N, NP, kDogLeg, false, true, false>(diag, L, st)
                  : launch_diag<double, N, NP, kDogLeg, false, false, false>(diag, L, st);
}

}  // namespace

}  // namespace gs
