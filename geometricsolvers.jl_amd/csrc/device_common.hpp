// Device-side scalar helpers of the batched solver: NaN-propagating max, the power-of-two scaling
// of the static LU (src/linear/factors.jl:726-919) and the norms of the reduction seam
// (src/base/reductions.jl:12-85), written for one lane that owns one system with every vector in
// registers (compile-time N, fully unrolled).
//
// The whole library is compiled with -ffp-contract=off: the reference's Julia arithmetic does not
// fuse `a*b + c`; the one place where it does (StaticArrays' `muladd` matrix-vector product) calls
// __builtin_fma explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

#define GS_DEV __device__ __forceinline__

constexpr double kDblMin = 2.2250738585072014e-308;     // floatmin(Float64)
constexpr double kDblEps = 2.220446049250313e-16;       // eps(Float64)
constexpr double kDblTiny = 4.9406564584124654e-324;    // nextfloat(0.0)
constexpr float kFltMin = 1.17549435e-38f;
constexpr int kEmax64 = 1023;                           // exponent(floatmax(Float64))

GS_DEV double dnan() { return __builtin_nan(""); }
GS_DEV bool isfin(double a) { return __builtin_isfinite(a); }
GS_DEV bool isnan_(double a) { return __builtin_isnan(a); }

// Julia's max/min of floats: NaN-propagating.
GS_DEV double jmax(double a, double b) { return (isnan_(a) || isnan_(b)) ? dnan() : (a > b ? a : b); }
GS_DEV double jmin(double a, double b) { return (isnan_(a) || isnan_(b)) ? dnan() : (a < b ? a : b); }

// a·2^e, exact unless it leaves the range (v_ldexp_f64 rounds correctly into the subnormals).
GS_DEV double ldexp2(double a, int e) { return __builtin_ldexp(a, e); }

// ⌊log2|a|⌋ of a finite nonzero a, subnormals included (Julia `exponent`).
GS_DEV int exponent_of(double a) { return __builtin_amdgcn_frexp_exp(a) - 1; }

// Squared 2-norm as a left fold without FMA (norm2, reductions.jl:12-15).
template <int L> GS_DEV double norm2(const double (&a)[L]) {
    double s = a[0] * a[0];
#pragma unroll
    for (int i = 1; i < L; ++i) s = s + a[i] * a[i];
    return s;
}

// The scaled path of rnorm (reductions.jl:355-363): it runs only for a sum that overflowed or fell
// below L·floatmin. Inline, so that no array becomes addressable (and no lane touches scratch).
template <int L> GS_DEV double rnorm_scaled(const double (&a)[L]) {
    double m = __builtin_fabs(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) m = jmax(m, __builtin_fabs(a[i]));
    if (m == 0.0 || !isfin(m)) return m;
    int e = exponent_of(m);
    double t = ldexp2(a[0], -e);
    double s = t * t;
#pragma unroll
    for (int i = 1; i < L; ++i) {
        t = ldexp2(a[i], -e);
        s = s + t * t;
    }
    return ldexp2(__builtin_sqrt(s), e);
}

// rnorm (reductions.jl:41-57).
template <int L> GS_DEV double rnorm(const double (&a)[L]) {
    double s = norm2(a);
    if (__builtin_expect(isfin(s) && s >= double(L) * kDblMin, 1)) return __builtin_sqrt(s);
    return rnorm_scaled(a);
}

// rmaxnorm (reductions.jl:67).
template <int L> GS_DEV double rmaxnorm(const double (&a)[L]) {
    double m = __builtin_fabs(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) m = jmax(m, __builtin_fabs(a[i]));
    return m;
}

// rdot (reductions.jl:77-85): a left fold of products, no FMA.
template <int L> GS_DEV double rdot(const double (&a)[L], const double (&b)[L]) {
    double s = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < L; ++i) s = s + a[i] * b[i];
    return s;
}

// The exponent of the largest |entry| of an array, clamped to ±emax; 0 when the array is zero or
// holds a non-finite entry (matrix_exponent, factors.jl:833-843, whose largest-abs-bits maximum makes
// a NaN the largest entry). A float max over |a| orders the finite entries as their bits do.
template <int L> GS_DEV int max_exponent(const double (&a)[L], int emax) {
    double m = __builtin_fabs(a[0]);
    bool nan_seen = isnan_(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) {
        m = __builtin_fmax(m, __builtin_fabs(a[i]));
        nan_seen |= isnan_(a[i]);
    }
    if (nan_seen || !isfin(m) || m == 0.0) return 0;
    int e = exponent_of(m);
    return e < -emax ? -emax : (e > emax ? emax : e);
}

// headroom(n) = 2 + ⌈log2 n⌉ (factors.jl:889-890), a compile-time constant.
constexpr int headroom(int n) {
    int b = 0;
    unsigned v = unsigned(n - 1);
    while (v) {
        ++b;
        v >>= 1;
    }
    return 2 + b;
}

// rhs_exponent (factors.jl:880-885) for a right-hand side of length L and factor exponent e.
template <int L> GS_DEV int rhs_exponent(const double (&r)[L], int e) {
    int m = max_exponent<L>(r, kEmax64);
    int lo = e < 0 ? e : 0;
    int c = m < lo ? lo : (m > 0 ? 0 : m);
    int alt = m - (kEmax64 - headroom(L));
    return c > alt ? c : alt;
}

}  // namespace gs
