// libgsolve_hip: the C ABI of include/gsolve.h over the MI355X kernels of this folder.
//
// gs_solve_batch_f64 replaces `_launch!` + `_solve_kernel!` of the reference's R2 batched solve
// (src/nonlinear/batched.jl:171-195): one launch over the batch, work-item i solving system
// idx = order[i] (or i), writing X[idx] and S[idx] (or codes[idx]), then a stream synchronisation.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/gsolve.h"
#include "mid_solver.hpp"
#include "small_launch.hpp"

namespace gs {

static thread_local std::string g_last_error;

// Record the message of a failed call for gs_last_error() and return its code.
int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

static int fail(int code, const std::string& msg) { return set_error(code, msg); }

// The kernels' view of a method and its options (options.jl:58-71; newton.jl:138-250).
DevParams dev_params(const gs_method& method, const gs_options_f64& options) {
    DevParams q;
    std::memset(&q, 0, sizeof(q));
    q.f_abstol = options.f_abstol;
    q.f_reltol = options.f_reltol;
    q.x_abstol = options.x_abstol;
    q.x_reltol = options.x_reltol;
    q.kappa = options.kappa;
    q.min_iterations = options.min_iterations;
    q.max_iterations = options.max_iterations;
    q.max_stalls = options.max_stalls;
    q.norm_max = options.norm == GS_NORM_MAX;
    q.roundoff = options.roundoff != 0;
    q.rate_test = options.rate_test != 0;
    q.alpha = method.static_alpha;
    q.c1 = method.bt_c1;
    q.p = method.bt_p;
    q.bt_maxiter = method.bt_maxiter;
    q.ref_maxiter = method.ref_maxiter;
    q.ref_eta_eff = method.ref_eta > 1.4901161193847656e-8 ? method.ref_eta : 1.4901161193847656e-8;
    q.mu = method.regularization;
    q.best = method.best != 0;
    q.reuse = method.reuse;
    q.dl_radius = method.dl_radius;
    return q;
}

// Kernel timing (gs_set_kernel_timing): events are created once per thread and reused.
static thread_local bool g_timing = false;
static thread_local double g_last_ms = -1.0;
static thread_local hipEvent_t g_ev[2] = {nullptr, nullptr};

// The bracket of gs_set_kernel_timing around one kernel launch on `st` (every entry point that
// times its kernel uses it): begin records the first event when timing is on.
bool timing_begin(hipStream_t st) {
    g_last_ms = -1.0;
    if (!g_timing) return false;
    if (!g_ev[0] && (hipEventCreate(&g_ev[0]) != hipSuccess || hipEventCreate(&g_ev[1]) != hipSuccess)) return false;
    return hipEventRecord(g_ev[0], st) == hipSuccess;
}

// Record the second event after the launch (before the caller synchronises).
bool timing_mark(hipStream_t st, bool timed) { return timed && hipEventRecord(g_ev[1], st) == hipSuccess; }

// After the synchronisation: the elapsed time for gs_last_kernel_ms.
void timing_end(bool timed) {
    float ms = 0.0f;
    g_last_ms = (timed && hipEventElapsedTime(&ms, g_ev[0], g_ev[1]) == hipSuccess) ? double(ms) : -1.0;
}

struct FamilySpec {
    int id, n, p_len, p_bytes;  // p_bytes: bytes of one record
    size_t align;               // alignment the device loads assume
};

static const FamilySpec* family_spec(int id) {
    static const FamilySpec specs[] = {
        {GS_FAM_COUPLED2, 2, 2, 16, 8},
        {GS_FAM_HETERO2, 2, 1, 8, 8},
        {GS_FAM_SCALAR_SQUARE, 1, 1, 8, 8},
        {GS_FAM_TANH_LINEAR3, 3, 12, 96, 8},
        {GS_FAM_HENON_HEILES_MIDPOINT, 4, 4, 32, 16},
        {GS_FAM_FPU_GL3, 48, 32, 256, 8},
    };
    for (const auto& s : specs)
        if (s.id == id) return &s;
    return nullptr;
}

}  // namespace gs

using namespace gs;

extern "C" {

const char* gs_last_error(void) { return g_last_error.c_str(); }

int32_t gs_version(void) { return GS_ABI_VERSION; }

void gs_abi_sizes(int64_t out[4]) {
    out[0] = sizeof(gs_status_f64);
    out[1] = sizeof(gs_options_f64);
    out[2] = sizeof(gs_method);
    out[3] = sizeof(gs_family);
}

int gs_solve_batch_f64(const gs_method* method, const gs_family* family, int64_t nsys, double* X, const void* P,
                       int64_t p_stride, gs_status_f64* S, uint8_t* codes, const int64_t* order,
                       const gs_options_f64* options, int32_t workgroupsize, void* stream) {
    g_last_error.clear();
    if (!method || !family || !options) return fail(GS_EINVAL, "method, family and options must not be NULL");
    if (nsys < 0) return fail(GS_EINVAL, "nsys must be non-negative");
    if (workgroupsize < 0 || workgroupsize > 1024)
        return fail(GS_EINVAL, "workgroupsize must be 0 (default) or in 1:1024, got " + std::to_string(workgroupsize));
    if (!S && !codes) return fail(GS_EINVAL, "one of S and codes must be given");
    const FamilySpec* fs = family_spec(family->id);
    if (!fs) return fail(GS_EFAMILY, "problem family " + std::to_string(family->id) + " is not registered");
    if (family->n != fs->n || family->m != fs->n)
        return fail(GS_EINVAL, "family " + std::to_string(family->id) + " has n = m = " + std::to_string(fs->n));
    if (method->linsolve != GS_LU) return fail(GS_EFAMILY, "only LUFactorization has a device form");
    if (method->globalization != GS_STATIC && method->globalization != GS_BACKTRACKING &&
        method->globalization != GS_DOGLEG)
        return fail(GS_EFAMILY, "globalization " + std::to_string(method->globalization) +
                                    " has no lane-per-system form for this family");
    if (method->globalization == GS_DOGLEG) {
        // the DogLeg constructor's checks (dogleg.jl:44-60): a radius stored clamped, or 0 (nothing)
        const double r = method->dl_radius;
        if (!(r == 0.0 || (r >= 1e-6 && r <= 1e10)))
            return fail(GS_EINVAL, "the DogLeg radius must be 0 (nothing) or in [1e-6, 1e10]");
        if (method->reuse != 0) return fail(GS_EINVAL, "Newton with a DogLeg globalization takes no reuse");
    }
    if (family->id == GS_FAM_FPU_GL3 && method->globalization != GS_STATIC)
        return fail(GS_EFAMILY, "the wave-per-system solve of FpuGl3 takes the Static globalization");
    if (method->reuse < 0) return fail(GS_EINVAL, "reuse must be ≥ 1, or 0 for nothing");
    if (family->id == GS_FAM_FPU_GL3 && method->reuse > 0)
        return fail(GS_EFAMILY, "the wave-per-system solve of FpuGl3 takes no reuse");
    if (method->factor_precision != GS_FACTOR_WORKING && method->factor_precision != GS_FACTOR_F32)
        return fail(GS_EINVAL, "factor_precision must be GS_FACTOR_WORKING or GS_FACTOR_F32");
    if (method->refinement != GS_NO_REFINEMENT && method->refinement != GS_REFINEMENT)
        return fail(GS_EINVAL, "refinement must be GS_NO_REFINEMENT or GS_REFINEMENT");
    if (nsys == 0) return GS_OK;
    if (!X) return fail(GS_EINVAL, "X must not be NULL for a non-empty batch");
    if (!P) return fail(GS_EINVAL, "this family needs its parameter records P");
    if (p_stride != 0 && p_stride < fs->p_bytes)
        return fail(GS_EINVAL, "p_stride " + std::to_string(p_stride) + " is below the record size " + std::to_string(fs->p_bytes));
    if ((reinterpret_cast<uintptr_t>(P) % fs->align) != 0 || (p_stride % int64_t(fs->align)) != 0)
        return fail(GS_EINVAL, "P and p_stride must be " + std::to_string(fs->align) + "-byte aligned for this family");
    if ((reinterpret_cast<uintptr_t>(X) % 8) != 0) return fail(GS_EINVAL, "X must be 8-byte aligned");

    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.fam.P = static_cast<const char*>(P);
    a.fam.p_stride = p_stride;
    for (int k = 0; k < 8; ++k) a.fam.shared[k] = family->shared[k];
    a.nsys = nsys;
    a.X = X;
    a.S = S;
    a.codes = codes;
    a.order = order;
    DevParams& q = a.q;
    q = dev_params(*method, *options);

    const bool diag = S != nullptr;
    if (!diag && !codes) return fail(GS_EINVAL, "codes must not be NULL");
    if (workgroupsize > kSmallMaxWorkgroup)
        return fail(GS_EINVAL, "workgroupsize " + std::to_string(workgroupsize) +
                                   " exceeds the solve kernel's maximum of " + std::to_string(kSmallMaxWorkgroup) +
                                   " work-items");
    const int wg = workgroupsize > 0 ? workgroupsize : kSmallMaxWorkgroup;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool mixed = method->factor_precision == GS_FACTOR_F32;
    const bool refine = method->refinement == GS_REFINEMENT;
    const int glob = method->globalization;
    bool timed = timing_begin(st);
    int rc;
    if (family->id == GS_FAM_FPU_GL3) {
        // one wave per system (the workgroup size is the kernel's own: 64)
        MidLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.q = q;
        L.P = static_cast<const char*>(P);
        L.p_stride = p_stride;
        for (int k = 0; k < 8; ++k) L.shared[k] = family->shared[k];
        L.nsys = nsys;
        L.X = X;
        L.S = S;
        L.codes = codes;
        L.order = order;
        L.mixed = mixed;
        L.refine = refine;
        rc = launch_fpu_gl3(L, st);
        if (rc != GS_OK) return rc;
        timed = timing_mark(st, timed);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(GS_EHIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return fail(GS_EHIP, std::string("hipStreamSynchronize failed: ") + hipGetErrorString(e));
        timing_end(timed);
        return GS_OK;
    }
    if (glob == GS_DOGLEG)
        rc = launch_small_dogleg(family->id, a, mixed, refine, diag, wg, st);
    else if (method->reuse > 0)
        rc = launch_small_reuse(family->id, glob, a, mixed, refine, diag, wg, st);
    else if (glob == GS_STATIC)
        rc = launch_small_family<kStatic, false>(family->id, a, mixed, refine, diag, wg, st);
    else
        rc = launch_small_family<kBacktracking, false>(family->id, a, mixed, refine, diag, wg, st);
    if (rc != GS_OK) return rc;
    timed = timing_mark(st, timed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GS_EHIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(GS_EHIP, std::string("hipStreamSynchronize failed: ") + hipGetErrorString(e));
    timing_end(timed);
    return GS_OK;
}

void gs_set_kernel_timing(int enable) { g_timing = enable != 0; }

double gs_last_kernel_ms(void) { return g_last_ms; }

}  // extern "C"
