/*
 * gsolve.h — C ABI of libgsolve_hip, the MI355X (gfx950) batched nonlinear solver.
 *
 * The library replaces the body of the R2 batched solve of GeometricSolvers.jl:
 *   solve_batch!(X, S, prob, method, p; order, options, workgroupsize)   src/nonlinear/batched.jl:257-287
 *   solve_batch!(X, prob, method, p; ...)  (counts form)                src/nonlinear/batched.jl:289-307
 * whose device part is the KernelAbstractions kernel `_solve_kernel!` (batched.jl:171-179) launched by
 * `_launch!` (batched.jl:183-195). A Julia package extension (ext/GeometricSolversHIPExt.jl) keeps the
 * Julia signature and `ccall`s the entry points below; the Python package in this repository binds
 * the same symbols through ctypes.
 *
 * Conventions
 *  - Every pointer to batch data (X, P, S, codes, order) is a device pointer of the caller.
 *  - Every call is synchronous: it enqueues on `stream` and waits for it before returning, as
 *    `_launch!` synchronises (batched.jl:193).
 *  - The return value is 0 or a negative GS_E* code; gs_last_error() then holds the message.
 *    Numerical failures never surface here: they are ReturnCodes inside S / codes.
 *  - No C++ exception crosses this boundary.
 *  - The library keeps device buffers of its own (workspaces, flags); calls that use them on one
 *    device are serialised by a per-device lock held from the first enqueue to the synchronisation,
 *    so two threads may call on the same device on different streams.
 */
#ifndef GSOLVE_H
#define GSOLVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 2

/* ReturnCode, one byte (src/base/returncodes.jl:14). */
enum gs_return_code {
    GS_SUCCESS = 0,
    GS_STALLED = 1,
    GS_MAXITERS = 2,
    GS_SINGULAR = 3,
    GS_NONFINITE = 4,
    GS_LINESEARCH_FAILED = 5
};

/* Error codes of the entry points (negative). */
enum gs_error {
    GS_OK = 0,
    GS_EINVAL = -1,     /* an argument is invalid: Julia ArgumentError                    */
    GS_EDIM = -2,       /* a length does not match: Julia DimensionMismatch               */
    GS_EFAMILY = -3,    /* the problem family or method combination is not registered     */
    GS_EHIP = -4        /* a HIP runtime call failed                                       */
};

/* SolverStatus{Float64} (src/base/status.jl:101-117), byte-identical to Julia's C layout:
 * code u8 @0, f_calls i16 @2, iterations i32 @4, fnorm @8, stepnorm @16, step_failures i32 @24,
 * promoted @28, stationary @29, jvp_calls i16 @30, rank i32 @32, nullnorm @40, gradratio @48,
 * rate @56, contracted @64, reused @65; size 72. */
typedef struct gs_status_f64 {
    uint8_t code;
    uint8_t pad0_;
    int16_t f_calls;
    int32_t iterations;
    double fnorm;
    double stepnorm;
    int32_t step_failures;
    uint8_t promoted;
    uint8_t stationary;
    int16_t jvp_calls;
    int32_t rank;
    int32_t pad1_;
    double nullnorm;
    double gradratio;
    double rate;
    uint8_t contracted;
    uint8_t reused;
    uint8_t pad2_[6];
} gs_status_f64;

/* SolverStatus{Float32}, byte-identical to Julia's C layout: code u8 @0, f_calls i16 @2, iterations
 * i32 @4, fnorm @8, stepnorm @12, step_failures i32 @16, promoted @20, stationary @21, jvp_calls i16
 * @22, rank i32 @24, nullnorm @28, gradratio @32, rate @36, contracted @40, reused @41; size 44. */
typedef struct gs_status_f32 {
    uint8_t code;
    uint8_t pad0_;
    int16_t f_calls;
    int32_t iterations;
    float fnorm;
    float stepnorm;
    int32_t step_failures;
    uint8_t promoted;
    uint8_t stationary;
    int16_t jvp_calls;
    int32_t rank;
    float nullnorm;
    float gradratio;
    float rate;
    uint8_t contracted;
    uint8_t reused;
    uint8_t pad1_[2];
} gs_status_f32;

/* The stopping norm of Options (src/base/options.jl:8-25). */
enum gs_norm { GS_NORM_L2 = 0, GS_NORM_MAX = 1 };

/* Options{Float64} (src/base/options.jl:58-71); `norm` is the StoppingNorm singleton as an enum. */
typedef struct gs_options_f64 {
    double f_abstol;
    double f_reltol;
    double x_abstol;
    double x_reltol;
    int32_t min_iterations;
    int32_t max_iterations;
    int32_t max_stalls;
    int32_t norm; /* gs_norm */
    double g_tol;
    double kappa;
    uint8_t roundoff;
    uint8_t rate_test;
    uint8_t pad_[6];
} gs_options_f64;

/* The method: Newton(; linsolve, globalization, refinement, regularization, best, reuse)
 * (src/nonlinear/newton.jl:138-250) with LUFactorization{Nothing|Float32} (src/linear/methods.jl:37-47),
 * Static (src/globalization/static.jl:8-30), Backtracking (src/globalization/backtracking.jl:28-50),
 * LevenbergMarquardt (src/globalization/levenbergmarquardt.jl:48-98) and Refinement
 * (src/nonlinear/refinement.jl:27-49), and DogLeg (src/globalization/dogleg.jl:1-80, the trust-region
 * step of newton.jl:1725-1780). `reuse = k` (newton.jl:1781-1874) is the period of the held
 * factor of the lane-per-system, n = 48 and dense solves with Static or Backtracking. `keep_factor`
 * keeps a factor across the solves of one state; a batch solves each system once, so it has no
 * field here. */
enum gs_globalization { GS_STATIC = 0, GS_BACKTRACKING = 1, GS_LEVENBERG_MARQUARDT = 2, GS_DOGLEG = 3 };
enum gs_factor_precision { GS_FACTOR_WORKING = 0 /* LUFactorization() */, GS_FACTOR_F32 = 1 };
enum gs_refinement { GS_NO_REFINEMENT = 0, GS_REFINEMENT = 1 };
enum gs_linsolve { GS_LU = 0 };

typedef struct gs_method {
    int32_t linsolve;          /* gs_linsolve                                            */
    int32_t factor_precision;  /* gs_factor_precision                                    */
    int32_t globalization;     /* gs_globalization                                       */
    int32_t refinement;        /* gs_refinement                                          */
    double static_alpha;       /* Static(α)                                              */
    double bt_c1;              /* Backtracking(c₁, p, maxiter)                           */
    double bt_p;
    int32_t bt_maxiter;
    int32_t ref_maxiter;       /* Refinement(η, maxiter)                                 */
    double ref_eta;
    double regularization;     /* μ ≥ 0                                                  */
    double lm_mu0;             /* LevenbergMarquardt(μ₀, μ_min, μ_max, μ_dec, μ_inc, ρ)  */
    double lm_mu_min;
    double lm_mu_max;
    double lm_mu_decrease;
    double lm_mu_increase;
    double lm_rho_accept;
    int32_t best;              /* best = true                                            */
    int32_t reuse;             /* reuse = k ≥ 1; 0 = nothing (factorise at every step)   */
    double dl_radius;          /* DogLeg(; radius) in [1e-6, 1e10]; 0 = nothing          */
    int32_t lm_scaling;        /* LevenbergMarquardt(; scaling): gs_lm_scaling           */
    int32_t pad_;
} gs_method;

/* LevenbergMarquardt(; scaling) (src/globalization/levenbergmarquardt.jl:318-347): the damping
 * diagonal D of [J; √μ D] δ = [−F; 0] is I, or Marquardt's column norms of J floored at
 * √eps · their largest. */
enum gs_lm_scaling { GS_LM_SCALING_NONE = 0, GS_LM_SCALING_DIAGONAL = 1 };

/* The registered problem families. Each has its residual F(x, p) and its analytic Jacobian J(x, p)
 * as HIP device functions; the Julia extension carries the same F and J as Julia functions, which
 * are the CPU oracle. `P` holds one record of `p_len` scalars per system at a byte stride of
 * `p_stride` (0 = one record shared by every system, as a `p` that is not a PerSystem). */
enum gs_family_id {
    GS_FAM_COUPLED2 = 1,         /* x + y²/8 − p : n = 2, P = 2 doubles (test/nonlinear/batched.jl:92-105)        */
    GS_FAM_HETERO2 = 2,          /* six-outcome fixture : n = 2, P = 1 int64 code (batched.jl:54-90)             */
    GS_FAM_SCALAR_SQUARE = 3,    /* x² − p : n = 1, P = 1 double (batched.jl:107-110)                            */
    GS_FAM_TANH_LINEAR3 = 4,     /* A_p x + 0.1 tanh(x) − b_p : n = 3, P = A (9, column-major) + b (3)           */
    GS_FAM_HENON_HEILES_MIDPOINT = 5, /* implicit-midpoint step of Hénon–Heiles : n = 4, P = x_k (4), shared[0] = h */
    GS_FAM_FPU_GL3 = 6,          /* Gauss–Legendre s = 3 stage system of a 16-site FPU-β chain : n = 48           */
    GS_FAM_EXPSUM_LSQ = 7,       /* exponential-sum least-squares fit (LevenbergMarquardt)                       */
    GS_ACT_HENON_HEILES = 8,     /* an action S(x, p), not a residual: the two-step discrete action of the       *
                                  * variational midpoint rule for Hénon–Heiles, x = (q₋, q, q₊) (6 entries),     *
                                  * P = (h, λ); taken by gs_solve_batch_stationarity_f64 only                   */
    GS_FAM_QUAD_MAP2 = 9,        /* u = A x, F_i = (u_i + c_i u_i²) − b_i : n = 2, P = A (4, column-major), b (2),  *
                                  * c (2) — f_mod_r1 / f_ill_r1 of test/nonlinear/refinement.jl:40-96           */
    GS_FAM_DENSE_SIN = 10        /* A x + sin.(x)/10 − b, the §10.E problem of benchmark/device_problems.jl:133-175 : *
                                  * n = 2, 4, 6, 8 (one lane per system) or 12, 16, 24, 32, 48, 64 (a segment    *
                                  * of 16, 32 or 64 lanes of a wave per system, gs_dense_set_shape); `data` = A  *
                                  * (n × n, column-major, shared by the batch, in the batch's precision), P = b  *
                                  * (n entries per system)                                                      */
};

typedef struct gs_family {
    int32_t id;       /* gs_family_id                                   */
    int32_t n;        /* unknowns per system                            */
    int32_t m;        /* residual entries per system (= n when square)  */
    int32_t p_len;    /* scalars per parameter record                   */
    double shared[8]; /* scalar parameters that every system shares     */
    const void* data; /* the family's own device data, which the reference keeps inside the problem
                       * (GS_FAM_DENSE_SIN: the matrix A of TenEResidual(A)); NULL for the others */
} gs_family;

/* Status form of solve_batch! (batched.jl:257-287, 311-321) on FP64 systems.
 *   X        nsys × n doubles, AoS (Julia Vector{SVector{n,Float64}}); overwritten with the iterates.
 *   P        parameter records (see gs_family); may be NULL for a family with p_len == 0.
 *   p_stride byte stride between records; 0 = shared record.
 *   S        nsys statuses (may be NULL when codes != NULL).
 *   codes    nsys ReturnCode bytes (may be NULL when S != NULL) — the counts form's buffer.
 *   order    NULL, or a permutation of 1..nsys (1-based, as Julia): work-item i solves order[i].
 *            The library checks it on the device before any solve (isperm, batched.jl:134-135):
 *            an entry outside 1..nsys or a repeated one returns GS_EINVAL with nothing written.
 *   workgroupsize 0 = library default, else the workgroup size of the launch (1..1024); a size
 *            above a kernel's own bound runs as workgroups of that bound (the systems are
 *            independent, so the grouping changes no result).
 * GS_FAM_FPU_GL3 and GS_FAM_DENSE_SIN take Static, Backtracking (with or without reuse) and DogLeg
 * (GS_EFAMILY otherwise).
 */
int gs_solve_batch_f64(const gs_method* method, const gs_family* family, int64_t nsys,
                       double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                       uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                       int32_t workgroupsize, void* stream);

/* gs_solve_batch_f64 with no allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, semantics and results as gs_solve_batch_f64. The device buffers the
 * call needs (the order check's bitmap, the FpuGl3 solve's hand-over flags) are a fixed array of
 * 2 MiB per GPU that the library holds until the process ends, so the call refuses (GS_EINVAL) an
 * `order` over more than 2^24 systems and an FpuGl3 batch of more than 2^20 systems. Every call on a
 * GPU shares that GPU's array: calls that overlapped on one GPU would race on it, and the library's
 * per-device lock keeps the calls of one process from overlapping. */
int gs_solve_batch_f64_noalloc(const gs_method* method, const gs_family* family, int64_t nsys,
                               double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                               uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                               int32_t workgroupsize, void* stream);

/* Free every device resource the library caches (the workspaces of the reductions, the
 * least-squares batch and the learned start, the least-squares graphs, streams and events, and the
 * calling thread's timing events) on every device, and restore the current device. No call may be
 * in flight; a later call creates what it needs again. No counterpart in the reference, whose KA
 * path (batched.jl:171-195) leaves device memory to Julia's garbage collector. */
int gs_shutdown(void);

/* Deflated(prob, roots; power, shift) (src/nonlinear/deflation.jl): the roots shared by every system
 * of a batch (the isbits Tuple of static roots that solve_batch! takes, deflation.jl:681-688). Root
 * r occupies roots[r·n .. r·n + n − 1]. */
typedef struct gs_deflation {
    int32_t nroots;     /* 1..4                                                               */
    int32_t pad_;
    double power;       /* p > 1/2, finite; the device takes integer-valued p (Julia's x^y then
                         * runs its integer power, Base.Math.pow_body)                       */
    double shift;       /* σ > 0, finite                                                      */
    double roots[32];   /* finite entries                                                     */
} gs_deflation;

/* gs_solve_batch_f64 on Deflated(NonlinearProblem(family), roots; power, shift): the steps and the
 * line search take the deflated residual M(x) F(x) with the product-rule Jacobian M J + F ∇Mᵀ
 * (deflation.jl:236-399), and the status reports, and the residual test reads, the inner
 * ‖F‖ = ‖M F‖ / M (converged_status / test_status, deflation.jl:514-548). Families 1–5, Static,
 * Backtracking and DogLeg, plain, mixed and refined, reuse = k under Static and Backtracking (the
 * refinement along a held factor takes the deflated JVP M (J v) + ⟨∇M, v⟩ F, deflation.jl:352-357);
 * a Newton with best = true is refused as the reference refuses it (deflation.jl:594-604). */
int gs_solve_batch_deflated_f64(const gs_method* method, const gs_family* family, const gs_deflation* deflation,
                                int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                int32_t workgroupsize, void* stream);

/* StationarityProblem(S; E, U, b) (src/nonlinear/stationarity.jl:1-60): the residual (∇S)_E + b on
 * the unknowns x[U] of each system's full state, the Jacobian (∇²S)_{E,U}. Indices are 1-based, as
 * the reference's; n_eq = 0 stands for E = U = : (every entry). */
enum gs_datum_kind {
    GS_DATUM_NONE = 0,          /* b = nothing                                                       */
    GS_DATUM_SHARED = 1,        /* one datum for every system: b_shared[0 .. |E|-1] (a Number repeated) */
    GS_DATUM_PER_SYSTEM = 2     /* PerSystem(b): system i reads |E| entries of the batch's precision at
                                   b_data + i·b_stride bytes (doubles; floats in the Float32 batch)   */
};

typedef struct gs_stationarity {
    int32_t n_eq;               /* |E| = |U|, or 0 for E = U = :                                      */
    int32_t b_kind;             /* gs_datum_kind                                                      */
    int32_t e[8];               /* E (1-based), n_eq entries                                          */
    int32_t u[8];               /* U (1-based), n_eq entries                                          */
    double b_shared[8];
    const void* b_data;         /* device                                                             */
    int64_t b_stride;           /* bytes between the datum records of two systems                     */
} gs_stationarity;

/* solve_batch! on a StationarityProblem (stationarity.jl:642-698, both forms): X holds each system's
 * full state (M doubles, M = 6 for GS_ACT_HENON_HEILES); the solve iterates the entries U and writes
 * them back, the others keep their bits; S or codes as for gs_solve_batch_f64. Served: the registered
 * actions, |E| = 2 or 6 (GS_EFAMILY otherwise), LUFactorization (Float64 or Float32 factor),
 * Static, Backtracking and DogLeg, refinement, regularization, best, reuse = k (not with DogLeg). */
int gs_solve_batch_stationarity_f64(const gs_method* method, const gs_family* action, const gs_stationarity* stat,
                                    int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                    uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                    int32_t workgroupsize, void* stream);

/* Reparametrized(NonlinearProblem(F), g) (src/nonlinear/reparametrized.jl:1-60, covariant = false):
 * Newton in chart coordinates y with F̃(y) = F(g(y)) and J̃ = J(g(y)) ∂g(y); the registered charts. */
enum gs_chart_kind {
    GS_CHART_LINEAR = 1,        /* x = B y + c, B n × n row-major in B[0 .. n²-1], c in c[0 .. n-1]   */
    GS_CHART_POLAR = 2          /* x = (y₁ cos y₂, y₁ sin y₂) (n = 2)                                  */
};

typedef struct gs_chart {
    int32_t kind;               /* gs_chart_kind                                                      */
    int32_t pad_;
    double B[16];
    double c[4];
} gs_chart;

/* solve_batch! on Reparametrized(NonlinearProblem(family), chart) (reparametrized.jl:981-1031, both
 * forms): X holds the chart coordinates y₀ on entry and receives x* = g(y*) (:860-865); S or codes
 * as for gs_solve_batch_f64. Families 1–5 with the linear chart, 1–2 with the polar chart; Static,
 * Backtracking and DogLeg, plain, mixed and refined, reuse = k (not with DogLeg; the JVP along a
 * held factor is J̃(y) v, reparametrized.jl:373-395). */
int gs_solve_batch_reparametrized_f64(const gs_method* method, const gs_family* family, const gs_chart* chart,
                                      int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                      uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                      int32_t workgroupsize, void* stream);

/* The Float32 twin of gs_solve_batch_f64: X holds nsys × n floats (Vector{SVector{n,Float32}}), the
 * parameter records hold floats (Int64 codes for family 2), S receives SolverStatus{Float32}. The
 * options and the method's floats arrive as doubles and are used as Float32 (R(x), as Options{Float32}
 * and ToReal{Float32} hold them; an Options{Float32} passes through exactly). A Float32 batch has no
 * mixed-precision form (its factor is in its working precision already: factor_precision is ignored).
 * The families of the lane-per-system kernels (1–5) and the dense family 10 (its matrix A in
 * gs_family.data as floats; Static, Backtracking with or without reuse, DogLeg) are served; family 6
 * is refused (GS_EFAMILY). */
int gs_solve_batch_f32(const gs_method* method, const gs_family* family, int64_t nsys,
                       float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                       uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                       int32_t workgroupsize, void* stream);

/* The Float32 twin of gs_solve_batch_stationarity_f64 (stationarity.jl:642-698 on a batch of
 * SVector{6,Float32} states): X holds nsys × M floats, the action's records (h, λ) are floats, a
 * shared datum is used as Float32 and a per-system datum holds floats (b_data, b_stride 4-byte
 * aligned); S receives SolverStatus{Float32}. Every form of the Float64 batch but the mixed one
 * (factor_precision is ignored, as in gs_solve_batch_f32). */
int gs_solve_batch_stationarity_f32(const gs_method* method, const gs_family* action, const gs_stationarity* stat,
                                    int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                    uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                    int32_t workgroupsize, void* stream);

/* The Float32 twin of gs_solve_batch_reparametrized_f64 (reparametrized.jl:981-1031 on a Float32
 * batch, the chart adapted by ToReal{Float32}): the linear chart's B and c are used as Float32, the
 * family's records hold floats as for gs_solve_batch_f32; S receives SolverStatus{Float32}. Families
 * 1–5 with the linear chart (the polar chart is refused with GS_EFAMILY), every form of the Float64
 * batch but the mixed one. */
int gs_solve_batch_reparametrized_f32(const gs_method* method, const gs_family* family, const gs_chart* chart,
                                      int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                      uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                      int32_t workgroupsize, void* stream);

/* The six-count tally of a codes buffer (batched.jl:197-199), on the device; counts[6] is host memory. */
int gs_tally_codes(const uint8_t* codes, int64_t nsys, int64_t counts[6], void* stream);

/* batch_summary(S) (batched.jl:80-89) on the device: out[0] = fnorm (the 2-norm of the per-system
 * fnorm), out_i[0] = failure_count, out_i[1] = max_iterations, out_i[2] = total_iterations. */
int gs_batch_summary_f64(const gs_status_f64* S, int64_t nsys, double* fnorm, int64_t out_i[3],
                         void* stream);

/* The least-squares batch: Newton(; globalization = LevenbergMarquardt()) on nfit exponential-sum
 * fits y(t) ≈ Σ_{k≤K} a_k exp(−λ_k t) of uneven sizes, all in one call.
 *
 * The reference solves a LeastSquaresProblem (src/nonlinear/leastsquares.jl:1-40) only one at a time,
 * `solve(prob, x₀, method, p)` (src/nonlinear/solver.jl:148-189, 277-378); its solve_batch! takes a
 * NonlinearProblem only (batched.jl:257-332). This entry point is that per-fit R1 solve for a batch,
 * with the status form's semantics (one SolverStatus per fit), and the LevenbergMarquardt step on the
 * stacked Givens reduction of [J | −F] (src/nonlinear/newton.jl:1876-1971;
 * src/globalization/levenbergmarquardt.jl:48-98, 162-164, 359-482, 575-584).
 *   X         nfit × 8 doubles; row i holds x = (a_1..a_K, λ_1..λ_K) in its first 2K slots; overwritten.
 *   K, m      int32 per fit: K ∈ 1..4 terms (n = 2K unknowns), m ∈ 2K..4096 points t_j = 5j/(m−1), j = 0..m−1.
 *   y_offset  int64 per fit: fit i's data is Y[y_offset[i] .. y_offset[i] + m[i] − 1], inside Y[0 .. ny − 1].
 *   S         nfit statuses.
 * Fits with K, m or y_offset out of range make the call fail with GS_EINVAL before any fit is solved. The fits are
 * bucketed by K on the device and run concurrently on library streams joined to `stream`. */
int gs_solve_expsum_lsq_f64(const gs_method* method, int64_t nfit, double* X, const int32_t* K,
                            const int32_t* m, const int64_t* y_offset, const double* Y, int64_t ny,
                            gs_status_f64* S, const gs_options_f64* options, void* stream);

/* The launch shape of the least-squares bucket K = k (1..4): `threads` per workgroup (64, 128, 192 or
 * 256) and resident workgroups per CU (`blocks_per_cu`, 1..64, or 0 for the kernel's occupancy) of its
 * persistent kernel (a lane whose fit ends takes the next fit of its bucket). Set by the package's
 * autotuner; no counterpart in the reference. gs_lsq_get_shape reports (threads, blocks_per_cu,
 * workgroups for a batch of nfit fits). */
int gs_lsq_set_shape(int k, int threads, int blocks_per_cu);
int gs_lsq_get_shape(int k, int64_t nfit, int64_t out[3]);

/* The launch shape of the dense family (GS_FAM_DENSE_SIN) at size n (12, 16, 24, 32, 48 or 64): `lanes`
 * per system, the wave segment that holds one system's rows (16, 32 or 64 and at least n; 16 lanes
 * put four systems in a wave, 64 one), or 0 for the default (the smallest that holds n). Results do
 * not depend on it. Set by the package's autotuner; no counterpart in the reference.
 * gs_dense_get_shape reports the lanes per system the next launch at size n uses (or a negative
 * error code). */
int gs_dense_set_shape(int n, int lanes);
int gs_dense_get_shape(int n);

/* The least-squares launch sequence (the device sort and the four bucket kernels) is captured into
 * a hipGraph and replayed while a call's pointers, sizes, method, options and shapes are those of the
 * capture (enable = 1, the default), or launched eagerly (0). No counterpart in the reference. */
int gs_lsq_set_graph(int enable);

/* The learned start, with no counterpart in the reference: a multilayer perceptron (inputs padded to
 * GS_MLP_IN, three hidden ReLU layers of GS_MLP_HIDDEN, outputs padded to GS_MLP_OUT; BF16 operands,
 * FP32 accumulation on MFMA) that predicts the Newton start of the implicit-midpoint step of
 * Hénon–Heiles (GS_FAM_HENON_HEILES_MIDPOINT) as a correction to the explicit midpoint step
 * x̃ = x + h f(x + (h/2) f(x)):
 *     x̂₊ = x̃ + h³ (o_j out_scale_j + out_shift_j),
 *     o = net(u),  u_j = (x_j − in_shift_j) in_scale_j (j < 4),  u_4 = (h − in_shift_4) in_scale_4.
 * params: GS_MLP_NPARAM floats (W1, b1, W2, b2, W3, b3, W4, b4; each W row-major, inputs × outputs);
 * packed: GS_MLP_NPACKED bfloat16 words, the weights in the MFMA operand order (gs_mlp_pack,
 * gs_mlp_adam write them). Newton in FP64 polishes the start; plain Newton from x stays the fallback. */
#define GS_MLP_IN 32
#define GS_MLP_HIDDEN 128
#define GS_MLP_OUT 16
#define GS_MLP_NPARAM 39312
#define GS_MLP_NPACKED 38912

typedef struct gs_mlp {
    const float* params;
    const uint16_t* packed;
    float in_shift[8];
    float in_scale[8];
    float out_shift[4];
    float out_scale[4];
} gs_mlp;

/* X0[i] = x̂₊ of Xk[i] (nsys × 4 doubles each, 16-byte aligned), on `stream` (asynchronous). */
int gs_mlp_start_henon_heiles(const gs_mlp* net, int64_t nsys, const double* Xk, double h, double* X0,
                              void* stream);

/* The fused learned solve: for each system, the start x̂₊ (as gs_mlp_start_henon_heiles), `method`
 * (Newton with LU, Static, FP64 or FP32 factors, with or without Refinement) from x̂₊, and plain
 * Newton from x₀ = x_k for every system whose solve from x̂₊ does not return SUCCESS; X and S receive
 * the final iterate and status of each system (the fallback's where it ran), *fallbacks the number of
 * systems that fell back. Synchronous. The reference has no learned start; per system this is its R1
 * solve!! (solver.jl:277-378) twice at most. */
int gs_solve_learned_henon_heiles_f64(const gs_mlp* net, const gs_method* method, int64_t nsys, const double* Xk,
                                      double h, double* X, gs_status_f64* S, const gs_options_f64* options,
                                      int64_t* fallbacks, void* stream);

/* Loss ½ mean_b,j<4 (o_j − t_j)² of a minibatch (systems idx[0..batch−1] of the training pairs
 * (Xk, X1), t the normalised correction of X1) into *loss (device float), and its gradient with respect
 * to every parameter into grad (GS_MLP_NPARAM device floats). batch is a multiple of 32. Asynchronous. */
int gs_mlp_grad(const gs_mlp* net, const double* Xk, const double* X1, double h, const int64_t* idx,
                int32_t batch, float* grad, float* loss, void* stream);

/* One Adam step (step ≥ 1, the bias corrections use it) on params with grad and the moments m, v
 * (all GS_MLP_NPARAM device floats), then the packed BF16 weights. Asynchronous. */
int gs_mlp_adam(float* params, uint16_t* packed, const float* grad, float* m, float* v, int64_t step, float lr,
                float beta1, float beta2, float eps, void* stream);

/* The packed BF16 weights of params. Asynchronous. */
int gs_mlp_pack(const float* params, uint16_t* packed, void* stream);

/* Instrumentation, with no counterpart in the reference: when enabled, gs_solve_batch_f64 and
 * gs_solve_learned_henon_heiles_f64 bracket their kernel with two HIP events on `stream`, and
 * gs_last_kernel_ms returns the elapsed time of the last such kernel of this thread (ms), or a
 * negative value if none was timed. */
void gs_set_kernel_timing(int enable);
double gs_last_kernel_ms(void);

/* The message of the last failed call on this thread ("" if none). */
const char* gs_last_error(void);

/* GS_ABI_VERSION, for the load-time check of a binding. */
int32_t gs_version(void);

/* sizeof(gs_status_f64), sizeof(gs_options_f64), sizeof(gs_method), sizeof(gs_family), in order:
 * a binding asserts these against its own layout at load time. */
void gs_abi_sizes(int64_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* GSOLVE_H */
