// libgsolve_hip, dense mid-size systems (n = 12, 16, 24, 32, 48, 64): the launch dispatch over the
// sizes and their lanes per system (the kernels are in dense_kernel.hpp, instantiated one size per
// translation unit in dense_n12.hip … dense_n64.hip, with reuse = k in dense_r12.hip …
// dense_r64.hip, under DogLeg in dense_d12.hip … dense_d64.hip, and in Float32 in dense_f16.hip …
// dense_f64.hip), and the shape setting of the autotuner.
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>

#include "dense_solver.hpp"

namespace gs {

int set_error(int code, const std::string& msg);

// The launches of one size, by lanes per system (dense_n16.hip … dense_n64.hip).
int launch_dense_n12(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_n16(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_n24(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_n32(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_n48(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_n64(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
// The Float32 launches (dense_f16.hip … dense_f64.hip).
int launch_dense_f32_n16(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st);
int launch_dense_f32_n24(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st);
int launch_dense_f32_n32(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st);
int launch_dense_f32_n48(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st);
int launch_dense_f32_n64(int np, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                         hipStream_t st);
// The DogLeg launches (dense_d12.hip … dense_d64.hip).
int launch_dense_dogleg_n12(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_dogleg_n16(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_dogleg_n24(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_dogleg_n32(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_dogleg_n48(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_dogleg_n64(int np, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
// The reuse = k launches (dense_r16.hip … dense_r64.hip).
int launch_dense_reuse_n12(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_reuse_n16(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_reuse_n24(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_reuse_n32(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_reuse_n48(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);
int launch_dense_reuse_n64(int np, int glob, bool mixed, bool refine, bool diag, const DenseLaunch& L, hipStream_t st);

namespace {

// The default lanes per system at size n: the smallest segment that holds the rows.
constexpr int default_np(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }

// The lanes per system of each size (0: the default), set by gs_dense_set_shape; index by size
// 12, 16, 24, 32, 48, 64.
std::atomic<int> g_dense_np[6] = {0, 0, 0, 0, 0, 0};
int size_slot(int n) {
    return n == 12 ? 0 : n == 16 ? 1 : n == 24 ? 2 : n == 32 ? 3 : n == 48 ? 4 : n == 64 ? 5 : -1;
}

// The lanes per system a size takes: 16 or 32 at n = 12 and 16, 32 or 64 at n = 24 and 32, 64 above.
bool lanes_allowed(int n, int lanes) {
    if (n <= 16) return lanes == 16 || lanes == 32;
    if (n == 24 || n == 32) return lanes == 32 || lanes == 64;
    return lanes == 64;
}

}  // namespace

int lanes_of(int n) {
    const int k = size_slot(n);
    const int np = k < 0 ? 0 : g_dense_np[k].load(std::memory_order_relaxed);
    return np != 0 ? np : default_np(n);
}

int dense_systems_per_block(int n) { return 64 / lanes_of(n) * 4; }

int dense_set_shape(int n, int lanes) {
    const int k = size_slot(n);
    if (k < 0) return set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
    if (lanes != 0 && !lanes_allowed(n, lanes))
        return set_error(GS_EINVAL, "lanes per system at n = " + std::to_string(n) + " must be 0 (default) or " +
                                        (n <= 16 ? "16 or 32" : (n <= 32 ? "32 or 64" : "64")));
    g_dense_np[k].store(lanes, std::memory_order_relaxed);
    return GS_OK;
}

int launch_dense(int n, int glob, bool mixed, bool refine, bool diag, bool reuse, const DenseLaunch& L,
                 hipStream_t st) {
    const int np = lanes_of(n);
    if (glob == kDogLeg) {
        if (reuse) return set_error(GS_EINVAL, "Newton with a DogLeg globalization takes no reuse");
        switch (n) {
            case 12: return launch_dense_dogleg_n12(np, mixed, refine, diag, L, st);
            case 16: return launch_dense_dogleg_n16(np, mixed, refine, diag, L, st);
            case 24: return launch_dense_dogleg_n24(np, mixed, refine, diag, L, st);
            case 32: return launch_dense_dogleg_n32(np, mixed, refine, diag, L, st);
            case 48: return launch_dense_dogleg_n48(np, mixed, refine, diag, L, st);
            case 64: return launch_dense_dogleg_n64(np, mixed, refine, diag, L, st);
        }
        return set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
    }
    if (reuse) {
        switch (n) {
            case 12: return launch_dense_reuse_n12(np, glob, mixed, refine, diag, L, st);
            case 16: return launch_dense_reuse_n16(np, glob, mixed, refine, diag, L, st);
            case 24: return launch_dense_reuse_n24(np, glob, mixed, refine, diag, L, st);
            case 32: return launch_dense_reuse_n32(np, glob, mixed, refine, diag, L, st);
            case 48: return launch_dense_reuse_n48(np, glob, mixed, refine, diag, L, st);
            case 64: return launch_dense_reuse_n64(np, glob, mixed, refine, diag, L, st);
        }
        return set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
    }
    switch (n) {
        case 12: return launch_dense_n12(np, glob, mixed, refine, diag, L, st);
        case 16: return launch_dense_n16(np, glob, mixed, refine, diag, L, st);
        case 24: return launch_dense_n24(np, glob, mixed, refine, diag, L, st);
        case 32: return launch_dense_n32(np, glob, mixed, refine, diag, L, st);
        case 48: return launch_dense_n48(np, glob, mixed, refine, diag, L, st);
        case 64: return launch_dense_n64(np, glob, mixed, refine, diag, L, st);
    }
    return set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
}

int launch_dense_f32(int n, int glob, bool refine, bool diag, bool reuse, const DenseLaunchT<float>& L,
                     hipStream_t st) {
    const int np = lanes_of(n);
    switch (n) {
        case 16: return launch_dense_f32_n16(np, glob, refine, diag, reuse, L, st);
        case 24: return launch_dense_f32_n24(np, glob, refine, diag, reuse, L, st);
        case 32: return launch_dense_f32_n32(np, glob, refine, diag, reuse, L, st);
        case 48: return launch_dense_f32_n48(np, glob, refine, diag, reuse, L, st);
        case 64: return launch_dense_f32_n64(np, glob, refine, diag, reuse, L, st);
    }
    return set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
}

}  // namespace gs

extern "C" {

int gs_dense_set_shape(int n, int lanes) { return gs::dense_set_shape(n, lanes); }

int gs_dense_get_shape(int n) {
    if (!(n == 12 || n == 16 || n == 24 || n == 32 || n == 48 || n == 64))
        return gs::set_error(GS_EINVAL, "the dense kernel takes n = 12, 16, 24, 32, 48 or 64");
    return gs::lanes_of(n);
}

}  // extern "C"
