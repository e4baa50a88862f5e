// Batch-level reductions of libgsolve_hip: the six-count tally of the counts form of solve_batch!
// (src/nonlinear/batched.jl:197-199, 289-307) and batch_summary (batched.jl:63-89).
//
// Each is a single pass of 256-thread workgroups, one partial per workgroup in a cached device
// workspace, and a fixed-order sum of the partials on the host, so that a result does not depend on
// the order in which workgroups finish.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gsolve.h"

namespace gs {

int set_error(int code, const std::string& msg);
// One call at a time per device, held through the synchronisation (defined with gs_solve_batch_f64).
std::unique_lock<std::recursive_mutex> device_call_lock();

namespace {

constexpr int kThreads = 256;

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

std::mutex g_ws_mutex;
Workspace g_ws[64];  // one per device ordinal

// A device workspace of at least `bytes`, kept until the process ends.
int workspace(size_t bytes, void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace& w = g_ws[dev];
    if (w.bytes < bytes) {
        if (w.ptr) (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
        if (hipMalloc(&w.ptr, bytes) != hipSuccess) return set_error(GS_EHIP, "hipMalloc of the reduction workspace failed");
        w.bytes = bytes;
    }
    *out = w.ptr;
    return GS_OK;
}

__global__ __launch_bounds__(kThreads) void tally_kernel(const uint8_t* codes, int64_t n, int64_t* partial) {
    __shared__ int counts[6];
    if (threadIdx.x < 6) counts[threadIdx.x] = 0;
    __syncthreads();
    int local[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        uint8_t c = codes[i];
#pragma unroll
        for (int k = 0; k < 6; ++k) local[k] += (c == k);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (local[k]) atomicAdd(&counts[k], local[k]);
    __syncthreads();
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = counts[threadIdx.x];
}

// Per workgroup: Σ fnorm², max |fnorm| (NaN-propagating), failures, max and Σ iterations.
struct SummaryPartial {
    double sumsq;
    double maxabs;
    int64_t failures;
    int64_t max_it;
    int64_t sum_it;
};

__device__ inline double nanmax(double a, double b) {
    return (__builtin_isnan(a) || __builtin_isnan(b)) ? __builtin_nan("") : (a > b ? a : b);
}

__global__ __launch_bounds__(kThreads) void summary_kernel(const gs_status_f64* S, int64_t n, int scale_e,
                                                           SummaryPartial* partial) {
    __shared__ double s_sum[kThreads];
    __shared__ double s_max[kThreads];
    __shared__ long long s_fail[kThreads], s_maxit[kThreads], s_sumit[kThreads];
    double sum = 0.0, mx = 0.0;
    long long fails = 0, maxit = 0, sumit = 0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        const gs_status_f64 s = S[i];
        double f = __builtin_ldexp(s.fnorm, -scale_e);
        sum = sum + f * f;
        mx = nanmax(mx, __builtin_fabs(s.fnorm));
        fails += s.code != GS_SUCCESS;
        maxit = s.iterations > maxit ? s.iterations : maxit;
        sumit += s.iterations;
    }
    s_sum[threadIdx.x] = sum;
    s_max[threadIdx.x] = mx;
    s_fail[threadIdx.x] = fails;
    s_maxit[threadIdx.x] = maxit;
    s_sumit[threadIdx.x] = sumit;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_sum[threadIdx.x] = s_sum[threadIdx.x] + s_sum[threadIdx.x + w];
            s_max[threadIdx.x] = nanmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
            s_fail[threadIdx.x] += s_fail[threadIdx.x + w];
            s_maxit[threadIdx.x] = s_maxit[threadIdx.x + w] > s_maxit[threadIdx.x] ? s_maxit[threadIdx.x + w] : s_maxit[threadIdx.x];
            s_sumit[threadIdx.x] += s_sumit[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = SummaryPartial{s_sum[0], s_max[0], s_fail[0], s_maxit[0], s_sumit[0]};
}

int blocks_for(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    return int(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

}  // namespace

// Free the workspace of every device (gs_shutdown; no call may be in flight).
void release_reduce() {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    for (int d = 0; d < 64; ++d) {
        if (!g_ws[d].ptr) continue;
        if (hipSetDevice(d) == hipSuccess) (void)hipFree(g_ws[d].ptr);
        g_ws[d] = Workspace{};
    }
}
}  // namespace gs

using namespace gs;

extern "C" {

int gs_tally_codes(const uint8_t* codes, int64_t nsys, int64_t counts[6], void* stream) {
    if (!counts) return set_error(GS_EINVAL, "counts must not be NULL");
    for (int k = 0; k < 6; ++k) counts[k] = 0;
    if (nsys < 0) return set_error(GS_EINVAL, "nsys must be non-negative");
    if (nsys == 0) return GS_OK;
    if (!codes) return set_error(GS_EINVAL, "codes must not be NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    const int blocks = blocks_for(nsys);
    void* ws = nullptr;
    int rc = workspace(size_t(blocks) * 6 * sizeof(int64_t), &ws);
    if (rc) return rc;
    hipLaunchKernelGGL(tally_kernel, dim3(blocks), dim3(kThreads), 0, st, codes, nsys, static_cast<int64_t*>(ws));
    std::vector<int64_t> host(size_t(blocks) * 6);
    if (hipMemcpyAsync(host.data(), ws, host.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return set_error(GS_EHIP, "tally copy failed");
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < 6; ++k) counts[k] += host[size_t(b) * 6 + k];
    return GS_OK;
}

int gs_batch_summary_f64(const gs_status_f64* S, int64_t nsys, double* fnorm, int64_t out_i[3], void* stream) {
    if (!fnorm || !out_i) return set_error(GS_EINVAL, "fnorm and out_i must not be NULL");
    *fnorm = 0.0;
    out_i[0] = out_i[1] = out_i[2] = 0;
    if (nsys < 0) return set_error(GS_EINVAL, "nsys must be non-negative");
    if (nsys == 0) return GS_OK;
    if (!S) return set_error(GS_EINVAL, "S must not be NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    const int blocks = blocks_for(nsys);
    void* ws = nullptr;
    int rc = workspace(size_t(blocks) * sizeof(SummaryPartial), &ws);
    if (rc) return rc;
    std::vector<SummaryPartial> host(blocks);
    auto pass = [&](int e) -> int {
        hipLaunchKernelGGL(summary_kernel, dim3(blocks), dim3(kThreads), 0, st, S, nsys, e, static_cast<SummaryPartial*>(ws));
        if (hipMemcpyAsync(host.data(), ws, host.size() * sizeof(SummaryPartial), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return set_error(GS_EHIP, "summary copy failed");
        return GS_OK;
    };
    if ((rc = pass(0))) return rc;
    double sum = 0.0, mx = 0.0;
    int64_t fails = 0, maxit = 0, sumit = 0;
    for (const auto& p : host) {
        sum = sum + p.sumsq;
        mx = (std::isnan(mx) || std::isnan(p.maxabs)) ? NAN : (p.maxabs > mx ? p.maxabs : mx);
        fails += p.failures;
        maxit = p.max_it > maxit ? p.max_it : maxit;
        sumit += p.sum_it;
    }
    // rnorm (src/base/reductions.jl:41-57): the scaled sum where the plain one overflows or underflows.
    double norm;
    if (std::isfinite(sum) && sum >= double(nsys) * 2.2250738585072014e-308) {
        norm = std::sqrt(sum);
    } else if (mx == 0.0 || !std::isfinite(mx)) {
        norm = mx;
    } else {
        int e = std::ilogb(mx);
        if ((rc = pass(e))) return rc;
        double s2 = 0.0;
        for (const auto& p : host) s2 = s2 + p.sumsq;
        norm = std::ldexp(std::sqrt(s2), e);
    }
    *fnorm = norm;
    out_i[0] = fails;
    out_i[1] = maxit;
    out_i[2] = sumit;
    return GS_OK;
}

}  // extern "C"
