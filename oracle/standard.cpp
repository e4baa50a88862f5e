// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// The 23 problems of the reference's standard test set (test/helpers/standard_problems.jl:55-386,
// after Moré, Garbow and Hillstrom 1981) restated as templates over the scalar type, with a
// forward-mode dual number for the Jacobian and the JVP, as the reference's cells take them from
// ForwardDiff (test/helpers/standard_cells.jl:128-148). `orc_standard_cell` runs one cell of the
// Newton/LU columns of the table with `max_iterations = 200` and default Options
// (standard_cells.jl:21, 33-35), so that tests/ can compare the oracle's solve!! loop with the
// golden table test/verification/standard_reference.csv (code, iterations, f_evals, fnorm).
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace orc {

using std::atan;
using std::cos;
using std::exp;
using std::sin;
using std::sqrt;

// A dual number with NP partials: value and derivative along each seeded direction.
template <class T, int NP> struct Dual {
    T v;
    T p[NP];
    Dual() : v(0) { for (int i = 0; i < NP; ++i) p[i] = 0; }
    Dual(T c) : v(c) { for (int i = 0; i < NP; ++i) p[i] = 0; }
};
template <class T, int NP> inline Dual<T, NP> operator+(Dual<T, NP> a, const Dual<T, NP>& b) {
    a.v = a.v + b.v;
    for (int i = 0; i < NP; ++i) a.p[i] = a.p[i] + b.p[i];
    return a;
}
template <class T, int NP> inline Dual<T, NP> operator-(Dual<T, NP> a, const Dual<T, NP>& b) {
    a.v = a.v - b.v;
    for (int i = 0; i < NP; ++i) a.p[i] = a.p[i] - b.p[i];
    return a;
}
template <class T, int NP> inline Dual<T, NP> operator-(Dual<T, NP> a) {
    a.v = -a.v;
    for (int i = 0; i < NP; ++i) a.p[i] = -a.p[i];
    return a;
}
template <class T, int NP> inline Dual<T, NP> operator*(const Dual<T, NP>& a, const Dual<T, NP>& b) {
    Dual<T, NP> r;
    r.v = a.v * b.v;
    for (int i = 0; i < NP; ++i) r.p[i] = a.p[i] * b.v + b.p[i] * a.v;
    return r;
}
template <class T, int NP> inline Dual<T, NP> operator/(const Dual<T, NP>& a, const Dual<T, NP>& b) {
    Dual<T, NP> r;
    r.v = a.v / b.v;
    T ib = T(1) / b.v;
    T c = -(a.v / (b.v * b.v));
    for (int i = 0; i < NP; ++i) r.p[i] = a.p[i] * ib + b.p[i] * c;
    return r;
}
template <class T, int NP> inline Dual<T, NP> chain(const Dual<T, NP>& a, T val, T der) {
    Dual<T, NP> r;
    r.v = val;
    for (int i = 0; i < NP; ++i) r.p[i] = der * a.p[i];
    return r;
}
template <class T, int NP> inline Dual<T, NP> sqrt(const Dual<T, NP>& a) {
    T s = std::sqrt(a.v);
    return chain(a, s, T(1) / (T(2) * s));
}
template <class T, int NP> inline Dual<T, NP> exp(const Dual<T, NP>& a) {
    T e = std::exp(a.v);
    return chain(a, e, e);
}
template <class T, int NP> inline Dual<T, NP> cos(const Dual<T, NP>& a) {
    return chain(a, std::cos(a.v), -std::sin(a.v));
}
template <class T, int NP> inline Dual<T, NP> sin(const Dual<T, NP>& a) {
    return chain(a, std::sin(a.v), std::cos(a.v));
}
template <class T, int NP> inline Dual<T, NP> atan(const Dual<T, NP>& a) {
    return chain(a, std::atan(a.v), T(1) / (T(1) + a.v * a.v));
}
template <class T> inline T cospi_(T x) { return std::cos(T(M_PI) * x); }
template <class T, int NP> inline Dual<T, NP> cospi_(const Dual<T, NP>& a) {
    return chain(a, std::cos(T(M_PI) * a.v), -T(M_PI) * std::sin(T(M_PI) * a.v));
}
template <class T> inline T sign_(T x) { return x > 0 ? T(1) : (x < 0 ? T(-1) : x); }
template <class T, int NP> inline Dual<T, NP> sign_(const Dual<T, NP>& a) { return Dual<T, NP>(sign_(a.v)); }
template <class T> inline T val(T x) { return x; }
template <class T, int NP> inline T val(const Dual<T, NP>& a) { return a.v; }
template <class S> inline S sq(const S& x) { return x * x; }
template <class S> inline S cube(const S& x) { return (x * x) * x; }
// x^2 and x^3 of a dual: ForwardDiff's rule for an integer power, `partials(x) * y * v^(y-1)`
// evaluated left to right (zero partials stay zero), with the value v^y of Base (v·v, (v·v)·v).
template <class T, int NP> inline Dual<T, NP> int_power(const Dual<T, NP>& a, T value, T y, T vpow) {
    Dual<T, NP> r(value);
    bool zero = true;
    for (int i = 0; i < NP; ++i) zero = zero && a.p[i] == T(0);
    if (!zero)
        for (int i = 0; i < NP; ++i) r.p[i] = (a.p[i] * y) * vpow;
    return r;
}
template <class T, int NP> inline Dual<T, NP> sq(const Dual<T, NP>& a) { return int_power(a, a.v * a.v, T(2), a.v); }
template <class T, int NP> inline Dual<T, NP> cube(const Dual<T, NP>& a) {
    return int_power(a, (a.v * a.v) * a.v, T(3), a.v * a.v);
}

// Scalar constants in the element type: C(c) converts a double literal, as T(c) in the reference.
#define C(c) S(T(c))

template <class T, class S> static void standard_F(int prob, int n, const S* x, S* out) {
    switch (prob) {
        case 1:
            out[0] = C(1) - x[0];
            for (int i = 1; i < n; ++i) out[i] = C(10) * (x[i] - x[i - 1] * x[i - 1]);
            break;
        case 2:
            out[0] = x[0] + C(10) * x[1];
            out[1] = C(std::sqrt(T(5))) * (x[2] - x[3]);
            out[2] = sq(x[1] - C(2) * x[2]);
            out[3] = (C(std::sqrt(T(10))) * (x[0] - x[3])) * (x[0] - x[3]);
            break;
        case 3:
            out[0] = (C(10000) * x[0]) * x[1] - C(1);
            out[1] = (exp(-x[0]) + exp(-x[1])) - C(1.0001);
            break;
        case 4: {
            S t1 = x[1] - x[0] * x[0];
            S t2 = x[3] - x[2] * x[2];
            out[0] = (C(-200) * x[0]) * t1 - (C(1) - x[0]);
            out[1] = (C(200) * t1 + C(20.2) * (x[1] - C(1))) + C(19.8) * (x[3] - C(1));
            out[2] = (C(-180) * x[2]) * t2 - (C(1) - x[2]);
            out[3] = (C(180) * t2 + C(20.2) * (x[3] - C(1))) + C(19.8) * (x[1] - C(1));
            break;
        }
        case 5: {
            S t1 = atan(x[1] / x[0]) / (C(2) * C(T(M_PI)));
            S t = val(x[0]) > T(0) ? t1 : (val(x[0]) < T(0) ? t1 + C(0.5) : C(0.25) * sign_(x[1]));
            out[0] = C(10) * (x[2] - C(10) * t);
            out[1] = C(10) * (sqrt(x[0] * x[0] + x[1] * x[1]) - C(1));
            out[2] = x[2];
            break;
        }
        case 6: {
            for (int k = 0; k < n; ++k) out[k] = C(0);
            for (int i = 1; i <= 29; ++i) {
                T ti = T(i) / T(29);
                S sum1 = C(0);
                T temp = T(1);
                for (int j = 2; j <= n; ++j) {
                    sum1 = sum1 + (C(T(j)) * C(temp)) * x[j - 1];
                    temp = ti * temp;
                }
                S sum2 = C(0);
                temp = T(1);
                for (int j = 1; j <= n; ++j) {
                    sum2 = sum2 + C(temp) * x[j - 1];
                    temp = ti * temp;
                }
                temp = T(1) / ti;
                for (int k = 1; k <= n; ++k) {
                    out[k - 1] = out[k - 1] + (C(temp) * ((sum1 - sum2 * sum2) - C(1))) *
                                                  (C(T(k)) - (C(T(2) * ti)) * sum2);
                    temp = ti * temp;
                }
            }
            out[0] = out[0] + x[0] * ((C(3) - C(2) * x[1]) + C(2) * sq(x[0]));
            out[1] = out[1] + (x[1] * (C(1) - x[1]) - C(1));
            break;
        }
        case 7: {
            for (int k = 0; k < n; ++k) out[k] = C(0);
            for (int j = 0; j < n; ++j) {
                S t1 = C(1);
                S t2 = x[j];
                for (int i = 0; i < n; ++i) {
                    out[i] = out[i] + t2;
                    S t3 = (C(2) * x[j]) * t2 - t1;
                    t1 = t2;
                    t2 = t3;
                }
            }
            for (int i = 1; i <= n; ++i) {
                T add = (i % 2 == 0) ? T(1) / T(i * i - 1) : T(0);
                out[i - 1] = out[i - 1] / C(T(n)) + C(add);
            }
            break;
        }
        case 8: {
            S sum = x[0];
            for (int i = 1; i < n; ++i) sum = sum + x[i];
            for (int i = 0; i < n - 1; ++i) out[i] = (x[i] + sum) - C(T(n + 1));
            S prod = x[0];
            for (int i = 1; i < n; ++i) prod = prod * x[i];
            out[n - 1] = prod - C(1);
            break;
        }
        case 9: {
            T h = T(1) / T(n + 1);
            auto ker = [&](const S& xk, int k) {
                return C(2) * xk + (C(T(0.5) * (h * h))) * cube((xk + C(T(k) * h)) + C(1));
            };
            out[0] = ker(x[0], 1) - x[1];
            for (int k = 2; k <= n - 1; ++k) out[k - 1] = (ker(x[k - 1], k) - x[k - 2]) - x[k];
            out[n - 1] = ker(x[n - 1], n) - x[n - 2];
            break;
        }
        case 10: {
            T h = T(1) / T(n + 1);
            for (int k = 1; k <= n; ++k) {
                T tk = T(k) / T(n + 1);
                S sum1 = C(0);
                for (int j = 1; j <= k; ++j) {
                    T tj = T(j) * h;
                    sum1 = sum1 + C(tj) * cube((x[j - 1] + C(tj)) + C(1));
                }
                S sum2 = C(0);
                for (int j = k; j <= n; ++j) {
                    T tj = T(j) * h;
                    sum2 = sum2 + C(T(1) - tj) * cube((x[j - 1] + C(tj)) + C(1));
                }
                out[k - 1] = x[k - 1] + (C(h) * (C(T(1) - tk) * sum1 + C(tk) * sum2)) / C(2);
            }
            break;
        }
        case 11: {
            S csum = cos(x[0]);
            for (int i = 1; i < n; ++i) csum = csum + cos(x[i]);
            for (int k = 1; k <= n; ++k) {
                S s = sin(x[k - 1]);
                S c = cos(x[k - 1]);
                out[k - 1] = ((C(T(n)) - csum) + C(T(k)) * (C(1) - c)) - s;
            }
            break;
        }
        case 12: {
            S sum1 = C(0);
            for (int j = 1; j <= n; ++j) sum1 = sum1 + C(T(j)) * (x[j - 1] - C(1));
            for (int k = 1; k <= n; ++k)
                out[k - 1] = (x[k - 1] - C(1)) + (C(T(k)) * sum1) * (C(1) + (C(2) * sum1) * sum1);
            break;
        }
        case 13: {
            auto ker = [&](const S& xk) { return (C(3) - C(2) * xk) * xk + C(1); };
            out[0] = ker(x[0]) - C(2) * x[1];
            for (int k = 2; k <= n - 1; ++k) out[k - 1] = (ker(x[k - 1]) - x[k - 2]) - C(2) * x[k];
            out[n - 1] = ker(x[n - 1]) - x[n - 2];
            break;
        }
        case 14: {
            for (int k = 1; k <= n; ++k) {
                int k1 = std::max(1, k - 5), k2 = std::min(n, k + 1);
                S temp = C(0);
                for (int j = k1; j <= k2; ++j)
                    if (j != k) temp = temp + x[j - 1] * (C(1) + x[j - 1]);
                out[k - 1] = (x[k - 1] * (C(2) + (C(5) * x[k - 1]) * x[k - 1]) + C(1)) - temp;
            }
            break;
        }
        case 15:
            out[0] = (x[0] * x[0] + x[1] * x[2]) - C(0.0001);
            out[1] = (x[0] * x[1] + x[1] * x[3]) - C(1);
            out[2] = (x[2] * x[0] + x[3] * x[2]) - C(0);
            out[3] = (x[2] * x[1] + x[3] * x[3]) - C(0.0001);
            break;
        case 16:
            out[0] = ((x[0] * x[0] + x[1] * x[3]) + x[2] * x[6]) - C(0.0001);
            out[1] = ((x[0] * x[1] + x[1] * x[4]) + x[2] * x[7]) - C(1);
            out[2] = (x[0] * x[2] + x[1] * x[5]) + x[2] * x[8];
            out[3] = (x[3] * x[0] + x[4] * x[3]) + x[5] * x[6];
            out[4] = ((x[3] * x[1] + x[4] * x[4]) + x[5] * x[7]) - C(0.0001);
            out[5] = (x[3] * x[2] + x[4] * x[5]) + x[5] * x[8];
            out[6] = (x[6] * x[0] + x[7] * x[3]) + x[8] * x[6];
            out[7] = (x[6] * x[1] + x[7] * x[4]) + x[8] * x[7];
            out[8] = ((x[6] * x[2] + x[7] * x[5]) + x[8] * x[8]) - C(0.0001);
            break;
        case 17:
            out[0] = (x[0] + x[1]) - C(3);
            out[1] = (sq(x[0]) + sq(x[1])) - C(9);
            break;
        case 18:
            out[0] = val(x[0]) == T(0) ? C(0) : (sq(x[1]) * (C(1) - exp((-x[0]) * x[0]))) / x[0];
            out[1] = val(x[1]) == T(0) ? C(0) : (x[0] * (C(1) - exp((-x[1]) * x[1]))) / x[1];
            break;
        case 19: {
            S t = sq(x[0]) + sq(x[1]);
            out[0] = x[0] * t;
            out[1] = x[1] * t;
            break;
        }
        case 20:
            out[0] = x[0] * sq(x[0] - C(5));
            break;
        case 21:
            out[0] = (((x[0] - cube(x[1])) + C(5) * sq(x[1])) - C(2) * x[1]) - C(13);
            out[1] = (((x[0] + cube(x[1])) + sq(x[1])) - C(14) * x[1]) - C(29);
            break;
        case 22:
            out[0] = (x[0] * x[0] - x[1]) + C(1);
            out[1] = x[0] - cospi_(C(0.5) * x[1]);
            break;
        case 23: {
            T c = T(0.9);
            for (int i = 0; i < n; ++i) out[i] = x[i];
            for (int i = 1; i <= n; ++i) {
                T mi = T(i) / T(n);
                S s = C(0);
                for (int j = 1; j <= n; ++j) {
                    T mj = T(j) / T(n);
                    s = s + (C(mi) * x[j - 1]) / C(mi + mj);
                }
                S term = C(1) - (C(c) * s) / C(T(2 * n));
                out[i - 1] = out[i - 1] - C(1) / term;
            }
            break;
        }
    }
}
#undef C

static int standard_n(int prob) {
    static const int ns[24] = {0, 10, 4, 2, 4, 3, 2, 2, 10, 10, 10, 10, 10, 10, 10, 4, 9, 2, 2, 2, 1, 2, 2, 10};
    return (prob >= 1 && prob <= 23) ? ns[prob] : 0;
}

// The start u0 of each problem in T (standard_problems.jl:392-423).
template <class T> static std::vector<T> standard_u0(int prob) {
    int n = standard_n(prob);
    std::vector<T> u(n);
    switch (prob) {
        case 1: u.assign(n, T(1)); u[0] = T(-1.2); break;
        case 2: u = {T(3), T(-1), T(0), T(1)}; break;
        case 3: u = {T(0), T(1)}; break;
        case 4: u = {T(-3), T(-1), T(-3), T(-1)}; break;
        case 5: u = {T(-1), T(0), T(0)}; break;
        case 6: u = {T(0), T(0)}; break;
        case 7: u = {T(0), T(2.0 / 3.0)}; break;
        case 8: u.assign(n, T(0.5)); break;
        case 9:
        case 10:
            for (int i = 1; i <= n; ++i) u[i - 1] = T(double(i * (i - 11)) / 121.0);
            break;
        case 11: u.assign(n, T(0.1)); break;
        case 12: for (int i = 1; i <= n; ++i) u[i - 1] = T(1.0 - double(i) / 10.0); break;
        case 13:
        case 14: u.assign(n, T(-1)); break;
        case 15: u = {T(1), T(0), T(0), T(1)}; break;
        case 16: u = {T(1), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(1)}; break;
        case 17: u = {T(1), T(5)}; break;
        case 18: u = {T(2), T(2)}; break;
        case 19: u = {T(3), T(3)}; break;
        case 20: u = {T(1)}; break;
        case 21: u = {T(0.5), T(-2)}; break;
        case 22: u = {T(1), T(0)}; break;
        case 23: u.assign(n, T(1)); break;
    }
    return u;
}

template <class T, int N> struct StandardProblem {
    static constexpr int NN = N;
    int prob;
    long nf = 0;
    Vec<T, N> F(const Vec<T, N>& x) {
        ++nf;
        Vec<T, N> out;
        standard_F<T, T>(prob, N, x.data(), out.data());
        return out;
    }
    // jvp!! with AutoForwardDiff (src/ad/jvp.jl:96-100): F at a dual iterate with one partial,
    // seeded along v; the partials of the residual are J(x) v. No primal F call is counted.
    Vec<T, N> Jv(const Vec<T, N>& x, const Vec<T, N>& v) const {
        using D = Dual<T, 1>;
        D xd[N], od[N];
        for (int j = 0; j < N; ++j) {
            xd[j] = D(x[j]);
            xd[j].p[0] = v[j];
        }
        standard_F<T, D>(prob, N, xd, od);
        Vec<T, N> out;
        for (int i = 0; i < N; ++i) out[i] = od[i].p[0];
        return out;
    }
    // The Jacobian by one forward pass of N partials (src/ad/forwarddiff_static.jl:50-55).
    Mat<T, N> J(const Vec<T, N>& x) const {
        using D = Dual<T, N>;
        D xd[N], od[N];
        for (int j = 0; j < N; ++j) {
            xd[j] = D(x[j]);
            xd[j].p[j] = T(1);
        }
        standard_F<T, D>(prob, N, xd, od);
        Mat<T, N> J;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) J[i + j * N] = od[i].p[j];
        return J;
    }
};

template <class T, int N>
static int run_cell(int prob, int scale, int column, int* code, int* iterations, long* f_evals,
                    double* fnorm, double* xout) {
    std::vector<T> u = standard_u0<T>(prob);
    Vec<T, N> x;
    for (int i = 0; i < N; ++i) x[i] = T(scale) * u[i];
    StandardProblem<T, N> p{prob};
    Method<T> m;
    // columns: 0 "Newton Static LU", 1 "Newton Backtracking LU", 2 "Newton reuse 3 Backtracking LU"
    // and 3 "Newton DogLeg LU" (standard_cells.jl:33-49)
    m.glob = column == 0 ? G_STATIC : column == 3 ? G_DOGLEG : G_BACKTRACKING;
    if (column == 2) m.reuse = 3;
    Options<T> o;
    o.max_iterations = 200;
    Solver<T, N, StandardProblem<T, N>> s(p, m, o);
    Status<T> st = s.solve(x);
    *code = st.code;
    *iterations = st.iterations;
    *f_evals = p.nf;
    *fnorm = double(st.fnorm);
    if (xout)
        for (int i = 0; i < N; ++i) xout[i] = double(x[i]);
    return 0;
}

template <class T>
static int dispatch_cell(int prob, int scale, int column, int* code, int* iterations, long* f_evals,
                         double* fnorm, double* xout) {
    switch (standard_n(prob)) {
        case 1: return run_cell<T, 1>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
        case 2: return run_cell<T, 2>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
        case 3: return run_cell<T, 3>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
        case 4: return run_cell<T, 4>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
        case 9: return run_cell<T, 9>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
        case 10: return run_cell<T, 10>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
    }
    return -1;
}

}  // namespace orc

extern "C" {

// One cell: problem 1..23, start scale 1/10/100, column 0 = "Newton Static LU",
// 1 = "Newton Backtracking LU", 2 = "Newton reuse 3 Backtracking LU", 3 = "Newton DogLeg LU",
// eltype 32 or 64. Writes the code, iterations, primal F calls and
// ‖F‖; xout (length n, may be NULL) receives the final iterate.
int orc_standard_cell(int prob, int scale, int column, int eltype, int* code, int* iterations,
                      long* f_evals, double* fnorm, double* xout) {
    if (eltype == 32)
        return orc::dispatch_cell<float>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
    return orc::dispatch_cell<double>(prob, scale, column, code, iterations, f_evals, fnorm, xout);
}

// The size n of a standard problem and its start u0 in Float64 (for the start-dedupe rule of
// standard_starts, standard_problems.jl:44-52).
int orc_standard_u0(int prob, int eltype, double* u) {
    int n = orc::standard_n(prob);
    if (u) {
        if (eltype == 32) {
            auto v = orc::standard_u0<float>(prob);
            for (int i = 0; i < n; ++i) u[i] = v[i];
        } else {
            auto v = orc::standard_u0<double>(prob);
            for (int i = 0; i < n; ++i) u[i] = v[i];
        }
    }
    return n;
}

}  // extern "C"
