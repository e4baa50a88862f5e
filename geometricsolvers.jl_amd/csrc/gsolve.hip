// libgsolve_hip: the C ABI of include/gsolve.h over the MI355X kernels of this folder.
//
// gs_solve_batch_f64 replaces `_launch!` + `_solve_kernel!` of the reference's R2 batched solve
// (src/nonlinear/batched.jl:171-195): one launch over the batch, work-item i solving system
// idx = order[i] (or i), writing X[idx] and S[idx] (or codes[idx]), then a stream synchronisation.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/gsolve.h"
#include "dense_solver.hpp"
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
template <class T> DevParamsT<T> dev_params_t(const gs_method& method, const gs_options_f64& options) {
    DevParamsT<T> q;
    std::memset(&q, 0, sizeof(q));
    q.f_abstol = T(options.f_abstol);
    q.f_reltol = T(options.f_reltol);
    q.x_abstol = T(options.x_abstol);
    q.x_reltol = T(options.x_reltol);
    q.kappa = T(options.kappa);
    q.min_iterations = options.min_iterations;
    q.max_iterations = options.max_iterations;
    q.max_stalls = options.max_stalls;
    q.norm_max = options.norm == GS_NORM_MAX;
    q.roundoff = options.roundoff != 0;
    q.rate_test = options.rate_test != 0;
    q.alpha = T(method.static_alpha);
    q.c1 = T(method.bt_c1);
    q.p = T(method.bt_p);
    q.bt_maxiter = method.bt_maxiter;
    q.ref_maxiter = method.ref_maxiter;
    const T eta = T(method.ref_eta);
    q.ref_eta_eff = eta > Lim<T>::sqrteps ? eta : Lim<T>::sqrteps;
    q.mu = T(method.regularization);
    q.best = method.best != 0;
    q.reuse = method.reuse;
    // DogLeg{R}(radius): the radius converted to R, then clamped to [1e-6, 1e10] in R
    T r = T(method.dl_radius);
    if (r != T(0)) r = r > T(1e10) ? T(1e10) : (r < T(1e-6) ? T(1e-6) : r);
    q.dl_radius = r;
    return q;
}

DevParams dev_params(const gs_method& method, const gs_options_f64& options) {
    return dev_params_t<double>(method, options);
}

// Kernel timing (gs_set_kernel_timing): events are created once per thread and reused.
// The release functions of the library's cached device resources (gs_shutdown).
void release_lsq();
void release_mlp();
void release_reduce();
void release_mid();

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
    int id, n, p_len;
    int p_bytes, align;          // one record and the alignment the device loads assume (Float64)
    int p_bytes32, align32;      // the same for a Float32 batch (0: no Float32 form)
};

// The sizes of the dense family: n ≤ 8 on the lane-per-system kernels, above on the dense kernel.
static bool dense_size(int n) {
    return n == 2 || n == 4 || n == 6 || n == 8 || n == 12 || n == 16 || n == 24 || n == 32 || n == 48 || n == 64;
}

static const FamilySpec* family_spec(int id) {
    static const FamilySpec specs[] = {
        {GS_FAM_COUPLED2, 2, 2, 16, 8, 8, 4},
        {GS_FAM_HETERO2, 2, 1, 8, 8, 8, 8},
        {GS_FAM_SCALAR_SQUARE, 1, 1, 8, 8, 4, 4},
        {GS_FAM_TANH_LINEAR3, 3, 12, 96, 8, 48, 4},
        {GS_FAM_QUAD_MAP2, 2, 8, 64, 8, 32, 4},
        {GS_FAM_HENON_HEILES_MIDPOINT, 4, 4, 32, 16, 16, 16},
        {GS_FAM_FPU_GL3, 48, 32, 256, 8, 0, 0},
        {GS_ACT_HENON_HEILES, 6, 2, 16, 8, 8, 4},
        {GS_FAM_DENSE_SIN, 0, 0, 0, 8, 0, 4},  // n = 16, 24, 32, 48 or 64, records of n entries
    };
    for (const auto& s : specs)
        if (s.id == id) return &s;
    return nullptr;
}

// The checks of both entry points, before any launch (batched.jl:113-150, the constructors' checks
// of the method); returns GS_OK or the error code with its message set.
template <class T>
static int validate_batch(const gs_method* method, const gs_family* family, int64_t nsys, const T* X, const void* P,
                          int64_t p_stride, const void* S, const uint8_t* codes, const gs_options_f64* options,
                          int32_t workgroupsize, const FamilySpec** spec) {
    constexpr bool f32 = sizeof(T) == 4;
    if (!method || !family || !options) return fail(GS_EINVAL, "method, family and options must not be NULL");
    if (nsys < 0) return fail(GS_EINVAL, "nsys must be non-negative");
    if (workgroupsize < 0 || workgroupsize > 1024)
        return fail(GS_EINVAL, "workgroupsize must be 0 (default) or in 1:1024, got " + std::to_string(workgroupsize));
    if (!S && !codes) return fail(GS_EINVAL, "one of S and codes must be given");
    const FamilySpec* fs = family_spec(family->id);
    if (!fs) return fail(GS_EFAMILY, "problem family " + std::to_string(family->id) + " is not registered");
    if (family->id == GS_FAM_DENSE_SIN) {
        if (!dense_size(family->n) || family->m != family->n)
            return fail(GS_EINVAL, "the dense family takes n = m = 2, 4, 6, 8, 12, 16, 24, 32, 48 or 64");
        if (!family->data || (reinterpret_cast<uintptr_t>(family->data) % 8) != 0)
            return fail(GS_EINVAL, "the dense family needs its matrix A (family data, 8-byte aligned)");
    } else if (family->n != fs->n || family->m != fs->n) {
        return fail(GS_EINVAL, "family " + std::to_string(family->id) + " has n = m = " + std::to_string(fs->n));
    }
    if (f32 && fs->p_bytes32 == 0 && family->id != GS_FAM_DENSE_SIN)
        return fail(GS_EFAMILY, "family " + std::to_string(family->id) + " has no Float32 form");
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
    if (method->reuse < 0) return fail(GS_EINVAL, "reuse must be ≥ 1, or 0 for nothing");
    if (method->factor_precision != GS_FACTOR_WORKING && method->factor_precision != GS_FACTOR_F32)
        return fail(GS_EINVAL, "factor_precision must be GS_FACTOR_WORKING or GS_FACTOR_F32");
    if (method->refinement != GS_NO_REFINEMENT && method->refinement != GS_REFINEMENT)
        return fail(GS_EINVAL, "refinement must be GS_NO_REFINEMENT or GS_REFINEMENT");
    *spec = fs;
    if (nsys == 0) return GS_OK;
    const int p_bytes = family->id == GS_FAM_DENSE_SIN ? int(sizeof(T)) * family->n : (f32 ? fs->p_bytes32 : fs->p_bytes);
    const int align = f32 ? fs->align32 : fs->align;
    if (!X) return fail(GS_EINVAL, "X must not be NULL for a non-empty batch");
    if (!P) return fail(GS_EINVAL, "this family needs its parameter records P");
    if (p_stride != 0 && p_stride < p_bytes)
        return fail(GS_EINVAL, "p_stride " + std::to_string(p_stride) + " is below the record size " + std::to_string(p_bytes));
    if ((reinterpret_cast<uintptr_t>(P) % align) != 0 || (p_stride % int64_t(align)) != 0)
        return fail(GS_EINVAL, "P and p_stride must be " + std::to_string(align) + "-byte aligned for this family");
    if ((reinterpret_cast<uintptr_t>(X) % sizeof(T)) != 0)
        return fail(GS_EINVAL, "X must be " + std::to_string(sizeof(T)) + "-byte aligned");
    return GS_OK;
}

// ---- One call at a time per device ----
// The library keeps device buffers of its own per device (workspaces, hand-over flags, the order
// check's bitmap, the no-allocation variant's `scratch` array). Every entry point that uses one holds
// its device's lock from its first enqueue through its synchronisation, so two calls on one device
// never share such a buffer at once, whatever streams they run on.
static std::recursive_mutex g_device_lock[64];

std::unique_lock<std::recursive_mutex> device_call_lock() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    return std::unique_lock<std::recursive_mutex>(g_device_lock[dev]);
}

// ---- isperm(order) on the device (batched.jl:133-135) ----
// Word 0 of `w` is the error flag, words 1.. a bitmap of the values seen; a fill kernel clears both
// (no memset node, so the sequence stays valid inside a captured graph).
__global__ __launch_bounds__(256) void order_clear_kernel(uint32_t* w, int64_t nwords) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < nwords; i += int64_t(gridDim.x) * 256) w[i] = 0u;
}

__global__ __launch_bounds__(256) void order_mark_kernel(const int64_t* order, int64_t nsys, uint32_t* w) {
    uint32_t* bits = w + 1;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < nsys; i += int64_t(gridDim.x) * 256) {
        const int64_t v = order[i];
        if (v < 1 || v > nsys) {
            atomicOr(w, 1u);  // out of 1:nsys
            continue;
        }
        const uint64_t k = uint64_t(v - 1);
        const uint32_t bit = 1u << (k & 31u);
        if (atomicOr(bits + (k >> 5), bit) & bit) atomicOr(w, 2u);  // a value repeats
    }
}

// Words of the check for a batch of nsys systems: the flag and ceil(nsys / 32) bitmap words.
static int64_t order_words(int64_t nsys) { return 1 + (nsys + 31) / 32; }

// The cached device buffer of the check (the entry points that allocate), per device; grown under
// the device's call lock, freed by gs_shutdown.
struct OrderWorkspace {
    uint32_t* ptr = nullptr;
    int64_t words = 0;
};
static OrderWorkspace g_order_ws[64];

static int order_workspace(int64_t words, uint32_t** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail(GS_EHIP, "hipGetDevice failed");
    OrderWorkspace& w = g_order_ws[dev];
    if (w.words < words) {
        if (w.ptr) (void)hipFree(w.ptr);
        w = OrderWorkspace{};
        if (hipMalloc(reinterpret_cast<void**>(&w.ptr), size_t(words) * 4) != hipSuccess)
            return fail(GS_EHIP, "hipMalloc of the order check's bitmap failed");
        w.words = words;
    }
    *out = w.ptr;
    return GS_OK;
}

void release_order() {
    for (int d = 0; d < 64; ++d) {
        if (!g_order_ws[d].ptr) continue;
        if (hipSetDevice(d) == hipSuccess) (void)hipFree(g_order_ws[d].ptr);
        g_order_ws[d] = OrderWorkspace{};
    }
}

// Refuse an `order` that is not a permutation of 1:nsys before any solve writes anything (the
// reference refuses it on the host, batched.jl:134-135). `w` is a device buffer of at least
// order_words(nsys) words, or null for the cached workspace. Synchronises `st`.
static int check_order(const int64_t* order, int64_t nsys, hipStream_t st, uint32_t* w) {
    if (!order || nsys == 0) return GS_OK;
    const int64_t words = order_words(nsys);
    if (!w) {
        const int rc = order_workspace(words, &w);
        if (rc != GS_OK) return rc;
    }
    const unsigned gc = unsigned(std::min<int64_t>((words + 255) / 256, 4096));
    const unsigned gm = unsigned(std::min<int64_t>((nsys + 255) / 256, 8192));
    hipLaunchKernelGGL(order_clear_kernel, dim3(gc), dim3(256), 0, st, w, words);
    hipLaunchKernelGGL(order_mark_kernel, dim3(gm), dim3(256), 0, st, order, nsys, w);
    uint32_t flag = 0;
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&flag, w, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(GS_EHIP, "the order check failed to run");
    if (flag != 0) return fail(GS_EINVAL, "order is not a permutation of 1:" + std::to_string(nsys));
    return GS_OK;
}

// ---- The no-allocation variant's scratch array ----
// One array per GPU the library runs on, held until the process ends. Within a call it serves, in
// turn, the order check (flag and bitmap: up to kNoAllocMaxOrdered systems) and then the n = 48
// solve's hand-over flags (one byte per system: up to kNoAllocMaxMid systems); the order check clears
// what it reads, and every n = 48 system writes its own flag before the second launch reads it.
constexpr int64_t kNoAllocMaxOrdered = int64_t(1) << 24;
constexpr int64_t kNoAllocMaxMid = int64_t(1) << 20;
constexpr size_t kScratchBytes = size_t(1 + (kNoAllocMaxOrdered + 31) / 32) * 4;
static_assert(kScratchBytes >= size_t(kNoAllocMaxMid), "the flags of the n = 48 solve fit the scratch array");
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static uint8_t* g_scratch_ptr[64] = {};

// The device address of this device's `scratch` (the array's name in host code is a stand-in, not a
// device address), looked up at the first call on each device, under the device's call lock.
static int scratch_address(uint8_t** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail(GS_EHIP, "hipGetDevice failed");
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GS_EHIP, std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<uint8_t*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return GS_OK;
}

// The workgroup of a lane-per-system launch: a requested size above the kernels' bound of
// kSmallMaxWorkgroup work-items runs as workgroups of that bound (the systems are independent, so
// how they group changes no result; batched.jl:200-210 leaves any size up to the device's maximum to
// the launch).
static int small_workgroup(int32_t w) { return (w <= 0 || w > kSmallMaxWorkgroup) ? kSmallMaxWorkgroup : w; }

// The launch, its timing bracket, and the synchronisation of `_launch!` (batched.jl:171-195).
static int finish_launch(int rc, hipStream_t st, bool timed) {
    if (rc != GS_OK) return rc;
    timed = timing_mark(st, timed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GS_EHIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(GS_EHIP, std::string("hipStreamSynchronize failed: ") + hipGetErrorString(e));
    timing_end(timed);
    return GS_OK;
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

}  // extern "C"

namespace gs {

// gs_solve_batch_f64, and with `noalloc` its variant that allocates nothing: the order check's
// buffer and the n = 48 solve's flags are then this device's `scratch` array.
static int solve_batch_f64(const gs_method* method, const gs_family* family, int64_t nsys, double* X, const void* P,
                           int64_t p_stride, gs_status_f64* S, uint8_t* codes, const int64_t* order,
                           const gs_options_f64* options, int32_t workgroupsize, void* stream, bool noalloc) {
    g_last_error.clear();
    const FamilySpec* fs = nullptr;
    int rc = validate_batch<double>(method, family, nsys, X, P, p_stride, S, codes, options, workgroupsize, &fs);
    if (rc != GS_OK) return rc;
    if (family->id == GS_ACT_HENON_HEILES)
        return set_error(GS_EFAMILY, "an action is solved as a StationarityProblem (gs_solve_batch_stationarity_f64)");
    if (nsys == 0) return rc;
    if (noalloc && order && nsys > kNoAllocMaxOrdered)
        return fail(GS_EINVAL, "gs_solve_batch_f64_noalloc checks an order of at most " +
                                   std::to_string(kNoAllocMaxOrdered) + " systems; got " + std::to_string(nsys));
    if (noalloc && family->id == GS_FAM_FPU_GL3 && nsys > kNoAllocMaxMid)
        return fail(GS_EINVAL, "gs_solve_batch_f64_noalloc solves at most " + std::to_string(kNoAllocMaxMid) +
                                   " FpuGl3 systems; got " + std::to_string(nsys));
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    uint8_t* scr = nullptr;
    if (noalloc && (rc = scratch_address(&scr)) != GS_OK) return rc;
    if ((rc = check_order(order, nsys, st, reinterpret_cast<uint32_t*>(scr))) != GS_OK) return rc;
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
    a.q = dev_params(*method, *options);
    const bool diag = S != nullptr;
    const int wg = small_workgroup(workgroupsize);
    const bool mixed = method->factor_precision == GS_FACTOR_F32;
    const bool refine = method->refinement == GS_REFINEMENT;
    const int glob = method->globalization;
    bool timed = timing_begin(st);
    if (family->id == GS_FAM_FPU_GL3) {
        // one wave per system (the workgroup size is the kernel's own: 64)
        MidLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.q = a.q;
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
        L.glob = glob == GS_BACKTRACKING ? kBacktracking : (glob == GS_DOGLEG ? kDogLeg : kStatic);
        L.reuse = method->reuse > 0;
        L.redo = scr;  // null: the library's own flags
        return finish_launch(launch_fpu_gl3(L, st), st, timed);
    }
    if (family->id == GS_FAM_DENSE_SIN && family->n <= 8) {
        // one lane per system, A read from device memory
        a.fam.data = family->data;
        const int g = glob == GS_BACKTRACKING ? kBacktracking : (glob == GS_DOGLEG ? kDogLeg : kStatic);
        return finish_launch(launch_small_dense(family->n, g, method->reuse > 0, a, mixed, refine, diag, wg, st), st,
                             timed);
    }
    if (family->id == GS_FAM_DENSE_SIN) {
        // a segment of a wave per system (the workgroup is the kernel's own: 256 work-items)
        DenseLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.q = a.q;
        L.A = static_cast<const double*>(family->data);
        L.P = static_cast<const char*>(P);
        L.p_stride = p_stride;
        L.nsys = nsys;
        L.X = X;
        L.S = S;
        L.codes = codes;
        L.order = order;
        const int g = glob == GS_BACKTRACKING ? kBacktracking : glob == GS_DOGLEG ? kDogLeg : kStatic;
        return finish_launch(launch_dense(family->n, g, mixed, refine, diag, method->reuse > 0, L, st), st, timed);
    }
    if (glob == GS_DOGLEG)
        rc = launch_small_dogleg(family->id, a, mixed, refine, diag, wg, st);
    else if (method->reuse > 0)
        rc = launch_small_reuse(family->id, glob, a, mixed, refine, diag, wg, st);
    else if (glob == GS_STATIC)
        rc = launch_small_family<kStatic, false>(family->id, a, mixed, refine, diag, wg, st);
    else
        rc = launch_small_family<kBacktracking, false>(family->id, a, mixed, refine, diag, wg, st);
    return finish_launch(rc, st, timed);
}

}  // namespace gs

// gs_solve_batch_stationarity_f64 and _f32: T is the working precision, Status its status record.
template <class T, class Status>
static int solve_stationarity(const gs_method* method, const gs_family* action, const gs_stationarity* stat,
                              int64_t nsys, T* X, const void* P, int64_t p_stride, Status* S, uint8_t* codes,
                              const int64_t* order, const gs_options_f64* options, int32_t workgroupsize,
                              void* stream) {
    g_last_error.clear();
    if (!stat) return set_error(GS_EINVAL, "stat must not be NULL");
    if (action && action->id != GS_ACT_HENON_HEILES)
        return set_error(GS_EFAMILY, "a StationarityProblem batch takes a registered action (GS_ACT_HENON_HEILES)");
    const FamilySpec* fs = nullptr;
    int rc = validate_batch<T>(method, action, nsys, X, P, p_stride, S, codes, options, workgroupsize, &fs);
    if (rc != GS_OK) return rc;
    // the constructor's index-set checks (stationarity.jl:87-103) and the solve's
    // (_validate_stationarity_solve, :133-157), then the device's own limits
    const gs_stationarity& t = *stat;
    const int M = fs->n;
    if (t.n_eq < 0 || t.n_eq > 8) return set_error(GS_EINVAL, "n_eq must be 0 (E = U = :) or in 1:8");
    const int ne = t.n_eq == 0 ? M : t.n_eq;
    for (int k = 0; k < t.n_eq; ++k) {
        if (t.e[k] < 1) return set_error(GS_EINVAL, "indices in E must be >= 1");
        if (t.u[k] < 1) return set_error(GS_EINVAL, "indices in U must be >= 1");
    }
    for (int k = 0; k < t.n_eq; ++k)
        for (int j = k + 1; j < t.n_eq; ++j) {
            if (t.e[k] == t.e[j]) return set_error(GS_EINVAL, "index set E has duplicate indices");
            if (t.u[k] == t.u[j]) return set_error(GS_EINVAL, "index set U has duplicate indices");
        }
    for (int k = 0; k < t.n_eq; ++k)
        if (t.e[k] > M) return set_error(GS_EINVAL, "index in E exceeds " + std::to_string(M));
    for (int k = 0; k < t.n_eq; ++k)
        if (t.u[k] > M) return set_error(GS_EINVAL, "index in U exceeds " + std::to_string(M));
    if (t.b_kind < GS_DATUM_NONE || t.b_kind > GS_DATUM_PER_SYSTEM) return set_error(GS_EINVAL, "b_kind must be a gs_datum_kind");
    if (ne != 2 && ne != 6)
        return set_error(GS_EFAMILY, "the device StationarityProblem batch takes |E| = 2 or 6 equations, got " +
                                         std::to_string(ne));
    if (nsys == 0) return GS_OK;
    if (t.b_kind == GS_DATUM_PER_SYSTEM) {
        if (!t.b_data) return set_error(GS_EINVAL, "a per-system datum needs b_data");
        constexpr int w = int(sizeof(T));
        if ((reinterpret_cast<uintptr_t>(t.b_data) % w) != 0 || (t.b_stride % w) != 0 || t.b_stride < w * ne)
            return set_error(GS_EINVAL, "b_data and b_stride must be " + std::to_string(w) +
                                            "-byte aligned, b_stride at least " + std::to_string(w) + "·|E| bytes");
    }
    KernelArgsT<T> a;
    std::memset(&a, 0, sizeof(a));
    a.fam.P = static_cast<const char*>(P);
    a.fam.p_stride = p_stride;
    for (int k = 0; k < 8; ++k) a.fam.shared[k] = action->shared[k];
    for (int k = 0; k < ne; ++k) {
        a.fam.stat.e[k] = t.n_eq == 0 ? k : t.e[k] - 1;
        a.fam.stat.u[k] = t.n_eq == 0 ? k : t.u[k] - 1;
        a.fam.stat.b_shared[k] = t.b_shared[k];
    }
    a.fam.stat.b_kind = t.b_kind;
    a.fam.stat.b_data = t.b_data;
    a.fam.stat.b_stride = t.b_stride;
    a.nsys = nsys;
    a.X = X;
    a.S = S;
    a.codes = codes;
    a.order = order;
    a.q = dev_params_t<T>(*method, *options);
    const bool diag = S != nullptr;
    const int wg = small_workgroup(workgroupsize);
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    if ((rc = check_order(order, nsys, st, nullptr)) != GS_OK) return rc;
    const bool mixed = method->factor_precision == GS_FACTOR_F32;
    const bool refine = method->refinement == GS_REFINEMENT;
    bool timed = timing_begin(st);
    const int glob = method->globalization == GS_BACKTRACKING ? kBacktracking
                     : method->globalization == GS_DOGLEG    ? kDogLeg
                                                             : kStatic;
    if constexpr (sizeof(T) == 8)
        rc = launch_small_stationarity(action->id, ne, glob, method->reuse > 0, a, mixed, refine, diag, wg, st);
    else
        rc = launch_small_stationarity_f32(action->id, ne, glob, method->reuse > 0, a, refine, diag, wg, st);
    return finish_launch(rc, st, timed);
}

// gs_solve_batch_reparametrized_f64 and _f32: T is the working precision, Status its status record.
template <class T, class Status>
static int solve_reparametrized(const gs_method* method, const gs_family* family, const gs_chart* chart, int64_t nsys,
                                T* X, const void* P, int64_t p_stride, Status* S, uint8_t* codes, const int64_t* order,
                                const gs_options_f64* options, int32_t workgroupsize, void* stream) {
    g_last_error.clear();
    const FamilySpec* fs = nullptr;
    int rc = validate_batch<T>(method, family, nsys, X, P, p_stride, S, codes, options, workgroupsize, &fs);
    if (rc != GS_OK) return rc;
    if (!chart) return set_error(GS_EINVAL, "chart must not be NULL");
    if (family->id == GS_FAM_FPU_GL3 || family->id == GS_ACT_HENON_HEILES)
        return set_error(GS_EFAMILY, "a Reparametrized batch wraps one of the lane-per-system families (1-5)");
    if (chart->kind == GS_CHART_POLAR) {
        if (sizeof(T) == 4) return set_error(GS_EFAMILY, "the Float32 Reparametrized batch takes the linear chart");
        if (fs->n != 2) return set_error(GS_EFAMILY, "the polar chart maps 2 coordinates: families 1 and 2");
    } else if (chart->kind == GS_CHART_LINEAR) {
        for (int k = 0; k < fs->n * fs->n; ++k)
            if (!std::isfinite(chart->B[k])) return set_error(GS_EINVAL, "the linear chart needs a finite B");
        for (int k = 0; k < fs->n; ++k)
            if (!std::isfinite(chart->c[k])) return set_error(GS_EINVAL, "the linear chart needs a finite c");
    } else {
        return set_error(GS_EINVAL, "chart kind must be GS_CHART_LINEAR or GS_CHART_POLAR");
    }
    if (nsys == 0) return GS_OK;
    KernelArgsT<T> a;
    std::memset(&a, 0, sizeof(a));
    a.fam.P = static_cast<const char*>(P);
    a.fam.p_stride = p_stride;
    for (int k = 0; k < 8; ++k) a.fam.shared[k] = family->shared[k];
    a.fam.chart.kind = chart->kind;
    for (int k = 0; k < 16; ++k) a.fam.chart.B[k] = chart->B[k];
    for (int k = 0; k < 4; ++k) a.fam.chart.c[k] = chart->c[k];
    a.nsys = nsys;
    a.X = X;
    a.S = S;
    a.codes = codes;
    a.order = order;
    a.q = dev_params_t<T>(*method, *options);
    const bool diag = S != nullptr;
    const int wg = small_workgroup(workgroupsize);
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    if ((rc = check_order(order, nsys, st, nullptr)) != GS_OK) return rc;
    const bool mixed = method->factor_precision == GS_FACTOR_F32;
    const bool refine = method->refinement == GS_REFINEMENT;
    bool timed = timing_begin(st);
    const int glob = method->globalization == GS_BACKTRACKING ? kBacktracking
                     : method->globalization == GS_DOGLEG    ? kDogLeg
                                                             : kStatic;
    if constexpr (sizeof(T) == 8)
        rc = launch_small_reparam(family->id, chart->kind, glob, method->reuse > 0, a, mixed, refine, diag, wg, st);
    else
        rc = launch_small_reparam_f32(family->id, chart->kind, glob, method->reuse > 0, a, refine, diag, wg, st);
    return finish_launch(rc, st, timed);
}

extern "C" {

int gs_solve_batch_f64(const gs_method* method, const gs_family* family, int64_t nsys, double* X, const void* P,
                       int64_t p_stride, gs_status_f64* S, uint8_t* codes, const int64_t* order,
                       const gs_options_f64* options, int32_t workgroupsize, void* stream) {
    return solve_batch_f64(method, family, nsys, X, P, p_stride, S, codes, order, options, workgroupsize, stream,
                           false);
}

int gs_solve_batch_f64_noalloc(const gs_method* method, const gs_family* family, int64_t nsys, double* X,
                               const void* P, int64_t p_stride, gs_status_f64* S, uint8_t* codes,
                               const int64_t* order, const gs_options_f64* options, int32_t workgroupsize,
                               void* stream) {
    return solve_batch_f64(method, family, nsys, X, P, p_stride, S, codes, order, options, workgroupsize, stream,
                           true);
}

int gs_solve_batch_deflated_f64(const gs_method* method, const gs_family* family, const gs_deflation* deflation,
                                int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                int32_t workgroupsize, void* stream) {
    g_last_error.clear();
    const FamilySpec* fs = nullptr;
    int rc = validate_batch<double>(method, family, nsys, X, P, p_stride, S, codes, options, workgroupsize, &fs);
    if (rc != GS_OK) return rc;
    // the Deflated constructor's checks (deflation.jl:120-132, 162-210) and _check_deflated_method
    // (:594-604), then the device's own limits
    if (!deflation) return set_error(GS_EINVAL, "deflation must not be NULL");
    const gs_deflation& d = *deflation;
    if (d.nroots < 1) return set_error(GS_EINVAL, "Deflated needs at least one root");
    if (d.nroots > 4) return set_error(GS_EINVAL, "the device deflation takes at most 4 roots");
    if (!(d.power > 0.5 && std::isfinite(d.power)))
        return set_error(GS_EINVAL,
                         "Deflated power must be finite and > 1/2: at or below 1/2 the known roots do not repel Newton");
    if (!(d.shift > 0.0 && std::isfinite(d.shift))) return set_error(GS_EINVAL, "Deflated shift must be finite and > 0");
    if (d.power != std::floor(d.power) || d.power > 64.0)
        return set_error(GS_EINVAL, "the device deflation takes an integer power in 1:64");
    for (int r = 0; r < d.nroots; ++r)
        for (int i = 0; i < fs->n; ++i)
            if (!std::isfinite(d.roots[r * fs->n + i]))
                return set_error(GS_EINVAL, "Deflated needs roots with finite real entries");
    if (method->best != 0)
        return set_error(GS_EINVAL,
                         "Deflated takes no Newton with best = true: the best iterate would be ranked by the deflated "
                         "residual, not by the inner residual that status.fnorm reports");
    if (family->id == GS_FAM_FPU_GL3 || family->id == GS_ACT_HENON_HEILES)
        return set_error(GS_EFAMILY, "a Deflated batch wraps one of the lane-per-system families (1-5)");
    if (nsys == 0) return GS_OK;
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.fam.P = static_cast<const char*>(P);
    a.fam.p_stride = p_stride;
    for (int k = 0; k < 8; ++k) a.fam.shared[k] = family->shared[k];
    a.fam.defl.nroots = d.nroots;
    a.fam.defl.power = int32_t(d.power);
    a.fam.defl.shift = d.shift;
    for (int k = 0; k < d.nroots * fs->n; ++k) a.fam.defl.roots[k] = d.roots[k];
    a.nsys = nsys;
    a.X = X;
    a.S = S;
    a.codes = codes;
    a.order = order;
    a.q = dev_params(*method, *options);
    const bool diag = S != nullptr;
    const int wg = small_workgroup(workgroupsize);
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    if ((rc = check_order(order, nsys, st, nullptr)) != GS_OK) return rc;
    const bool mixed = method->factor_precision == GS_FACTOR_F32;
    const bool refine = method->refinement == GS_REFINEMENT;
    bool timed = timing_begin(st);
    const int glob = method->globalization == GS_BACKTRACKING ? kBacktracking
                     : method->globalization == GS_DOGLEG    ? kDogLeg
                                                             : kStatic;
    rc = launch_small_deflated(family->id, glob, method->reuse > 0, a, mixed, refine, diag, wg, st);
    return finish_launch(rc, st, timed);
}

int gs_solve_batch_stationarity_f64(const gs_method* method, const gs_family* action, const gs_stationarity* stat,
                                    int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                    uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                    int32_t workgroupsize, void* stream) {
    return solve_stationarity(method, action, stat, nsys, X, P, p_stride, S, codes, order, options, workgroupsize,
                              stream);
}

int gs_solve_batch_stationarity_f32(const gs_method* method, const gs_family* action, const gs_stationarity* stat,
                                    int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                    uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                    int32_t workgroupsize, void* stream) {
    return solve_stationarity(method, action, stat, nsys, X, P, p_stride, S, codes, order, options, workgroupsize,
                              stream);
}

int gs_solve_batch_reparametrized_f64(const gs_method* method, const gs_family* family, const gs_chart* chart,
                                      int64_t nsys, double* X, const void* P, int64_t p_stride, gs_status_f64* S,
                                      uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                      int32_t workgroupsize, void* stream) {
    return solve_reparametrized(method, family, chart, nsys, X, P, p_stride, S, codes, order, options, workgroupsize,
                                stream);
}

int gs_solve_batch_reparametrized_f32(const gs_method* method, const gs_family* family, const gs_chart* chart,
                                      int64_t nsys, float* X, const void* P, int64_t p_stride, gs_status_f32* S,
                                      uint8_t* codes, const int64_t* order, const gs_options_f64* options,
                                      int32_t workgroupsize, void* stream) {
    return solve_reparametrized(method, family, chart, nsys, X, P, p_stride, S, codes, order, options, workgroupsize,
                                stream);
}

int gs_solve_batch_f32(const gs_method* method, const gs_family* family, int64_t nsys, float* X, const void* P,
                       int64_t p_stride, gs_status_f32* S, uint8_t* codes, const int64_t* order,
                       const gs_options_f64* options, int32_t workgroupsize, void* stream) {
    g_last_error.clear();
    const FamilySpec* fs = nullptr;
    int rc = validate_batch<float>(method, family, nsys, X, P, p_stride, S, codes, options, workgroupsize, &fs);
    if (rc != GS_OK) return rc;
    if (family->id == GS_ACT_HENON_HEILES)
        return set_error(GS_EFAMILY, "an action is solved as a StationarityProblem (gs_solve_batch_stationarity_f32)");
    if (nsys == 0) return rc;
    KernelArgsT<float> a;
    std::memset(&a, 0, sizeof(a));
    a.fam.P = static_cast<const char*>(P);
    a.fam.p_stride = p_stride;
    for (int k = 0; k < 8; ++k) a.fam.shared[k] = family->shared[k];
    a.nsys = nsys;
    a.X = X;
    a.S = S;
    a.codes = codes;
    a.order = order;
    a.q = dev_params_t<float>(*method, *options);
    const int wg = small_workgroup(workgroupsize);
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    if ((rc = check_order(order, nsys, st, nullptr)) != GS_OK) return rc;
    bool timed = timing_begin(st);
    if (family->id == GS_FAM_DENSE_SIN && family->n <= 12) {
        // one lane per system, A read from device memory (at n = 12 too: in Float32 the rows fit a lane,
        // measured 1.35e9 against 7.4e8 systems/s on the segment kernel for the grid's gs-newton cell)
        a.fam.data = family->data;
        const int g = method->globalization == GS_BACKTRACKING ? kBacktracking
                      : method->globalization == GS_DOGLEG    ? kDogLeg
                                                              : kStatic;
        rc = launch_small_dense_f32(family->n, g, method->reuse > 0, a, method->refinement == GS_REFINEMENT,
                                    S != nullptr, wg, st);
        return finish_launch(rc, st, timed);
    }
    if (family->id == GS_FAM_DENSE_SIN) {
        // a segment of a wave per system; the factor is in the working precision (Float32)
        DenseLaunchT<float> L;
        std::memset(&L, 0, sizeof(L));
        L.q = a.q;
        L.A = static_cast<const float*>(family->data);
        L.P = static_cast<const char*>(P);
        L.p_stride = p_stride;
        L.nsys = nsys;
        L.X = X;
        L.S = S;
        L.codes = codes;
        L.order = order;
        const int g = method->globalization == GS_BACKTRACKING ? kBacktracking
                      : method->globalization == GS_DOGLEG    ? kDogLeg
                                                              : kStatic;
        rc = launch_dense_f32(family->n, g, method->refinement == GS_REFINEMENT, S != nullptr, method->reuse > 0, L, st);
        return finish_launch(rc, st, timed);
    }
    rc = launch_small_f32(family->id, method->globalization, method->reuse > 0, a,
                          method->refinement == GS_REFINEMENT, S != nullptr, wg, st);
    return finish_launch(rc, st, timed);
}

void gs_set_kernel_timing(int enable) { g_timing = enable != 0; }

int gs_shutdown(void) {
    int dev = 0;
    const bool have = hipGetDevice(&dev) == hipSuccess;
    release_lsq();
    release_mlp();
    release_reduce();
    release_mid();
    release_order();
    if (g_ev[0]) (void)hipEventDestroy(g_ev[0]);
    if (g_ev[1]) (void)hipEventDestroy(g_ev[1]);
    g_ev[0] = g_ev[1] = nullptr;
    if (have) (void)hipSetDevice(dev);
    return GS_OK;
}

double gs_last_kernel_ms(void) { return g_last_ms; }

}  // extern "C"
