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

// The constants of the working precision R (Float64 or Float32): floatmin, eps, nextfloat(0),
// exponent(floatmax) and √eps.
template <class T> struct Lim;
template <> struct Lim<double> {
    static constexpr double min = 2.2250738585072014e-308, eps = 2.220446049250313e-16, tiny = 4.9406564584124654e-324;
    static constexpr int emax = 1023;
    static constexpr double sqrteps = 1.4901161193847656e-8;
};
template <> struct Lim<float> {
    static constexpr float min = 1.17549435e-38f, eps = 1.1920929e-7f, tiny = 1.4e-45f;
    static constexpr int emax = 127;
    static constexpr float sqrteps = 3.4526698e-4f;
};

GS_DEV double dnan() { return __builtin_nan(""); }
template <class T> GS_DEV T tnan() { return T(__builtin_nan("")); }
GS_DEV bool isfin(double a) { return __builtin_isfinite(a); }
GS_DEV bool isfin(float a) { return __builtin_isfinite(a); }
GS_DEV bool isnan_(double a) { return __builtin_isnan(a); }
GS_DEV bool isnan_(float a) { return __builtin_isnan(a); }
GS_DEV double tabs(double a) { return __builtin_fabs(a); }
GS_DEV float tabs(float a) { return __builtin_fabsf(a); }
GS_DEV double tsqrt(double a) { return __builtin_sqrt(a); }
GS_DEV float tsqrt(float a) { return __builtin_sqrtf(a); }
GS_DEV double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
GS_DEV float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GS_DEV double tfmax(double a, double b) { return __builtin_fmax(a, b); }
GS_DEV float tfmax(float a, float b) { return __builtin_fmaxf(a, b); }
GS_DEV double tfmin(double a, double b) { return __builtin_fmin(a, b); }
GS_DEV float tfmin(float a, float b) { return __builtin_fminf(a, b); }

// Julia's max/min of floats: NaN-propagating.
template <class T> GS_DEV T jmax(T a, T b) { return (isnan_(a) || isnan_(b)) ? tnan<T>() : (a > b ? a : b); }
template <class T> GS_DEV T jmin(T a, T b) { return (isnan_(a) || isnan_(b)) ? tnan<T>() : (a < b ? a : b); }

// a·2^e, exact unless it leaves the range (v_ldexp rounds correctly into the subnormals).
GS_DEV double ldexp2(double a, int e) { return __builtin_ldexp(a, e); }
GS_DEV float ldexp2(float a, int e) { return __builtin_ldexpf(a, e); }

// ⌊log2|a|⌋ of a finite nonzero a, subnormals included (Julia `exponent`).
GS_DEV int exponent_of(double a) { return __builtin_amdgcn_frexp_exp(a) - 1; }
GS_DEV int exponent_of(float a) { return __builtin_amdgcn_frexp_expf(a) - 1; }

// Squared 2-norm as a left fold without FMA (norm2, reductions.jl:12-15).
template <int L, class T> GS_DEV T norm2(const T (&a)[L]) {
    T s = a[0] * a[0];
#pragma unroll
    for (int i = 1; i < L; ++i) s = s + a[i] * a[i];
    return s;
}

// The scaled path of rnorm (reductions.jl:355-363): it runs only for a sum that overflowed or fell
// below L·floatmin. Inline, so that no array becomes addressable (and no lane touches scratch).
template <int L, class T> GS_DEV T rnorm_scaled(const T (&a)[L]) {
    T m = tabs(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) m = jmax(m, tabs(a[i]));
    if (m == T(0) || !isfin(m)) return m;
    int e = exponent_of(m);
    T t = ldexp2(a[0], -e);
    T s = t * t;
#pragma unroll
    for (int i = 1; i < L; ++i) {
        t = ldexp2(a[i], -e);
        s = s + t * t;
    }
    return ldexp2(tsqrt(s), e);
}

// The correctly rounded square root of a finite s ≥ 2^-767: the operations LLVM's AMDGPU backend
// emits for an FP64 sqrt (a reciprocal square root refined by two Goldschmidt steps and two
// residual corrections, every step a fused multiply-add) without its range scaling (by 2^256 below
// 2^-767) and its select of s itself for ±0 and +∞, which are identities in that range. Same
// operands, same operations: the bits of __builtin_sqrt(s), in 10 instructions instead of 21.
GS_DEV double sqrt_inrange(double s) {
    double r = __builtin_amdgcn_rsq(s);
    double g = s * r;
    double h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    double d = __builtin_fma(-g, g, s);
    h = __builtin_fma(h, e, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, s);
    return __builtin_fma(d, h, g);
}
constexpr double kSqrtInrangeMin = 0x1p-767;

// rnorm (reductions.jl:41-57); the sum is in the element type (Float32 or Float64).
template <int L, class T> GS_DEV T rnorm(const T (&a)[L]) {
    T s = norm2<L>(a);
    if constexpr (sizeof(T) == 8) {
        if (__builtin_expect(isfin(s) && s >= kSqrtInrangeMin, 1)) return sqrt_inrange(s);
    }
    if (__builtin_expect(isfin(s) && s >= T(L) * Lim<T>::min, 1)) return tsqrt(s);
    return rnorm_scaled<L>(a);
}

// The largest |entry| of an array as a fold of IEEE maxNum, which passes over NaN entries (NaN only
// when every entry is NaN).
template <int L, class T> GS_DEV T max_abs(const T (&a)[L]) {
    T m = tabs(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) m = tfmax(m, tabs(a[i]));
    return m;
}

// rmaxnorm (reductions.jl:67).
template <int L, class T> GS_DEV T rmaxnorm(const T (&a)[L]) {
    T m = tabs(a[0]);
#pragma unroll
    for (int i = 1; i < L; ++i) m = jmax(m, tabs(a[i]));
    return m;
}

// rdot (reductions.jl:77-85): a left fold of products, no FMA.
template <int L, class T> GS_DEV T rdot(const T (&a)[L], const T (&b)[L]) {
    T s = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < L; ++i) s = s + a[i] * b[i];
    return s;
}

// The exponent of the largest |entry| of an array, clamped to ±emax; 0 when the array is zero or
// holds a non-finite entry (matrix_exponent, factors.jl:833-843, whose largest-abs-bits maximum makes
// a NaN the largest entry). A float max over |a| orders the finite entries as their bits do.
template <int L, class T> GS_DEV int max_exponent(const T (&a)[L], int emax) {
    const T m = max_abs<L>(a);
    bool nan_seen = false;
#pragma unroll
    for (int i = 0; i < L; ++i) nan_seen |= isnan_(a[i]);
    if (nan_seen || !isfin(m) || m == T(0)) return 0;
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

// 2^e as a compile-time constant of type T.
template <class T> constexpr T exp2i(int e) {
    T v = T(1);
    for (int i = 0; i < e; ++i) v *= T(2);
    for (int i = 0; i > e; --i) v /= T(2);
    return v;
}

// rhs_exponent (factors.jl:880-885) for a right-hand side of length L and factor exponent e.
template <int L, class T> GS_DEV int rhs_exponent(const T (&r)[L], int e) {
    constexpr int emax = Lim<T>::emax;
    int m = max_exponent<L>(r, emax);
    int lo = e < 0 ? e : 0;
    int c = m < lo ? lo : (m > 0 ? 0 : m);
    int alt = m - (emax - headroom(L));
    return c > alt ? c : alt;
}

}  // namespace gs
