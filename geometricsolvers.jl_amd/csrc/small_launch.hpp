// The lane-per-system kernel of gs_solve_batch_f64 and its launch, one globalization at a time, so
// that the instantiations of each form (plain, reuse = k, DogLeg) compile in translation units of
// their own: gsolve.hip (Static and Backtracking), small_reuse.hip (reuse = k), small_dogleg.hip
// (DogLeg).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "../../include/gsolve.h"
#include "families.hpp"
#include "small_solver.hpp"

namespace gs {

int set_error(int code, const std::string& msg);

template <class T> struct KernelArgsT {
    FamArgs fam;
    DevParamsT<T> q;
    int64_t nsys;
    T* X;
    void* S;  // gs_status_f64 or gs_status_f32
    uint8_t* codes;
    const int64_t* order;
    int32_t stage = 0;  // 1: rows and statuses of whole waves go through LDS (set by launch_diag)
};
using KernelArgs = KernelArgsT<double>;

// The status of one system in the byte layout of SolverStatus{Float64}, as nine 8-byte words.
GS_DEV void pack_status(uint64_t (&w)[9], uint8_t code, int32_t iterations, double fnorm, double stepnorm,
                        int32_t failures, bool promoted, double gradratio, double rate, bool contracted, bool reused) {
    w[0] = uint64_t(code) | (uint64_t(uint32_t(iterations)) << 32);
    w[1] = __builtin_bit_cast(uint64_t, fnorm);
    w[2] = __builtin_bit_cast(uint64_t, stepnorm);
    w[3] = uint64_t(uint32_t(failures)) | (uint64_t(promoted ? 1 : 0) << 32);
    w[4] = uint64_t(uint32_t(-1));  // rank = -1: an LU factor reports no rank
    w[5] = __builtin_bit_cast(uint64_t, dnan());  // nullnorm
    w[6] = __builtin_bit_cast(uint64_t, gradratio);
    w[7] = __builtin_bit_cast(uint64_t, rate);
    w[8] = uint64_t(contracted ? 1 : 0) | (uint64_t(reused ? 1 : 0) << 8);
}

// The status of one system in the byte layout of SolverStatus{Float32} (44 bytes), as eleven 4-byte
// words: code @0, f_calls @2, iterations @4, fnorm @8, stepnorm @12, step_failures @16, promoted @20,
// stationary @21, jvp_calls @22, rank @24, nullnorm @28, gradratio @32, rate @36, contracted @40,
// reused @41.
GS_DEV void pack_status(uint32_t (&w)[11], uint8_t code, int32_t iterations, float fnorm, float stepnorm,
                        int32_t failures, bool promoted, float gradratio, float rate, bool contracted, bool reused) {
    w[0] = uint32_t(code);
    w[1] = uint32_t(iterations);
    w[2] = __builtin_bit_cast(uint32_t, fnorm);
    w[3] = __builtin_bit_cast(uint32_t, stepnorm);
    w[4] = uint32_t(failures);
    w[5] = uint32_t(promoted ? 1 : 0);
    w[6] = uint32_t(-1);  // rank = -1: an LU factor reports no rank
    w[7] = __builtin_bit_cast(uint32_t, tnan<float>());  // nullnorm
    w[8] = __builtin_bit_cast(uint32_t, gradratio);
    w[9] = __builtin_bit_cast(uint32_t, rate);
    w[10] = uint32_t(contracted ? 1 : 0) | (uint32_t(reused ? 1 : 0) << 8);
}

// The status words of one working precision: 9 × 8 bytes (Float64) or 11 × 4 bytes (Float32).
template <class T> struct StatusWords {
    using W = uint64_t;
    using Status = gs_status_f64;
    static constexpr int kCount = 9;
};
template <> struct StatusWords<float> {
    using W = uint32_t;
    using Status = gs_status_f32;
    static constexpr int kCount = 11;
};

template <class W, int NW> GS_DEV void store_status_lane(void* dst, const W (&w)[NW]) {
    W* d = static_cast<W*>(dst);
#pragma unroll
    for (int k = 0; k < NW; ++k) d[k] = w[k];
}

GS_DEV void store_status(gs_status_f64* dst, uint8_t code, int32_t iterations, double fnorm, double stepnorm,
                         int32_t failures, bool promoted, double gradratio, double rate, bool contracted,
                         bool reused) {
    uint64_t w[9];
    pack_status(w, code, iterations, fnorm, stepnorm, failures, promoted, gradratio, rate, contracted, reused);
    store_status_lane(dst, w);
}

// LDS bytes per wave of the staged status store (64 statuses).
template <class T> constexpr int status_stage_bytes() {
    return 64 * StatusWords<T>::kCount * int(sizeof(typename StatusWords<T>::W));
}

// LDS bytes per wave of the staged stores of a lane-per-system kernel: the statuses, then the rows
// of X, one after the other in the same block.
template <class T, int N> constexpr int small_stage_bytes() {
    return status_stage_bytes<T>() > 64 * N * int(sizeof(T)) ? status_stage_bytes<T>() : 64 * N * int(sizeof(T));
}

// The statuses of a wave's 64 consecutive systems, written as one contiguous span of 64 × 72 (or 44)
// bytes: each lane puts its words in the wave's LDS block, and the wave then stores the block in
// 16-byte pieces of consecutive lanes, so that every store instruction covers whole cache lines
// (a lane's own 72-byte record, stored word by word, makes each of the nine store instructions touch
// all 36 lines of the span). `lds` and `dst` are 16-byte aligned.
// The rows of X of a wave's 64 consecutive systems (N entries each) as one contiguous span, through
// LDS like store_status_wave; `lds` and `dst` are 16-byte aligned.
template <class T, int N> GS_DEV void store_rows_wave(T* lds, T* dst, const T (&v)[N]) {
    const int lane = int(threadIdx.x & 63u);
#pragma unroll
    for (int k = 0; k < N; ++k) lds[lane * N + k] = v[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int kPieces = 64 * N * int(sizeof(T)) / 16;
    const uint4* src = reinterpret_cast<const uint4*>(lds);
    uint4* out = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int r = 0; r < (kPieces + 63) / 64; ++r) {
        const int c = lane + 64 * r;
        if (kPieces % 64 == 0 || c < kPieces) out[c] = src[c];
    }
    // the block's reads are done before the wave writes it again
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class W, int NW> GS_DEV void store_status_wave(W* lds, void* dst, const W (&w)[NW]) {
    const int lane = int(threadIdx.x & 63u);
#pragma unroll
    for (int k = 0; k < NW; ++k) lds[lane * NW + k] = w[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int kPieces = 64 * NW * int(sizeof(W)) / 16;
    const uint4* src = reinterpret_cast<const uint4*>(lds);
    uint4* out = static_cast<uint4*>(dst);
#pragma unroll
    for (int r = 0; r < (kPieces + 63) / 64; ++r) {
        const int c = lane + 64 * r;
        if (kPieces % 64 == 0 || c < kPieces) out[c] = src[c];
    }
}

// The largest workgroup of the lane-per-system kernels (their launch bound) and their default.
constexpr int kSmallMaxWorkgroup = 256;

// Waves per SIMD the register allocator targets: 4 (≤ 128 VGPRs) for the plain FP64 Static form of
// the residual families, which fit without scratch; the heavier forms, and the families that hold a
// full state (a StationarityProblem's action), keep the allocator's own choice.
template <class Fam, int GLOB, bool MIXED, bool REFINE, bool REUSE> constexpr int small_waves_per_eu() {
    return (GLOB == kStatic && !MIXED && !REFINE && !REUSE && !holds_state<Fam>::value) ? 4 : 1;
}

// One work-item per system (the launch shape of `_solve_kernel!`, batched.jl:171-179).
template <class Fam, int GLOB, bool MIXED, bool REFINE, bool DIAG, bool REUSE>
__global__ __launch_bounds__(kSmallMaxWorkgroup)
__attribute__((amdgpu_waves_per_eu(small_waves_per_eu<Fam, GLOB, MIXED, REFINE, REUSE>())))
void solve_small_kernel(const KernelArgsT<typename Fam::T> a) {
    using T = typename Fam::T;
    using Status = typename std::conditional<sizeof(T) == 8, gs_status_f64, gs_status_f32>::type;
    constexpr int N = Fam::N;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    // A wave of 64 consecutive systems in natural order writes its rows of X and its statuses through
    // LDS (store_rows_wave, store_status_wave); the launch sizes the LDS for it exactly when X and the
    // status buffer are 16-byte aligned, the batch has no `order` and the workgroup is whole waves
    // (a.stage).
    const bool staged = DIAG && a.stage && int64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u) + 64 <= a.nsys;
    if (i >= a.nsys) return;
    const int64_t idx = a.order ? a.order[i] - 1 : i;
    Fam fam;
    fam.load(a.fam, idx);
    T x[N];
    if constexpr (holds_state<Fam>::value) {
        // a row of X is the full state; the unknowns are its entries U
        T* xp = a.X + idx * Fam::kStateLen;
        fam.gather(xp, x);
        SmallSolve<Fam, GLOB, MIXED, REFINE, DIAG, REUSE> solver(fam, a.q);
        auto R = solver.solve(x);
        fam.scatter(xp, x);
        if (DIAG) {
            typename StatusWords<T>::W w[StatusWords<T>::kCount];
            pack_status(w, R.code, R.iterations, R.fnorm, R.stepnorm, R.failures, R.promoted, R.gradratio, R.rate,
                        R.contracted, R.reused);
            store_status_lane(static_cast<Status*>(a.S) + idx, w);
        } else {
            a.codes[idx] = R.code;
        }
        return;
    }
    T* xp = a.X + idx * N;
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = xp[k];
    SmallSolve<Fam, GLOB, MIXED, REFINE, DIAG, REUSE> solver(fam, a.q);
    auto R = solver.solve(x);
    T xo[N];
    if constexpr (maps_output<Fam>::value) {
        // the batch receives x* = g(y*) of a Reparametrized problem's chart
        fam.output(x, xo);
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) xo[k] = x[k];
    }
    if (DIAG) {
        typename StatusWords<T>::W w[StatusWords<T>::kCount];
        pack_status(w, R.code, R.iterations, R.fnorm, R.stepnorm, R.failures, R.promoted, R.gradratio, R.rate,
                    R.contracted, R.reused);
        if (staged) {
            extern __shared__ __align__(16) unsigned char gs_status_stage[];
            unsigned char* blk = gs_status_stage + (threadIdx.x >> 6) * small_stage_bytes<T, N>();
            const int64_t i0 = i - int64_t(threadIdx.x & 63u);
            store_rows_wave(reinterpret_cast<T*>(blk), a.X + i0 * N, xo);
            store_status_wave(reinterpret_cast<typename StatusWords<T>::W*>(blk), static_cast<Status*>(a.S) + i0, w);
        } else {
#pragma unroll
            for (int k = 0; k < N; ++k) xp[k] = xo[k];
            store_status_lane(static_cast<Status*>(a.S) + idx, w);
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) xp[k] = xo[k];
        a.codes[idx] = R.code;
    }
}

template <class Fam, int GLOB, bool MIXED, bool REFINE, bool REUSE>
int launch_diag(const KernelArgsT<typename Fam::T>& a, bool diag, int wg, hipStream_t st) {
    const int64_t blocks = (a.nsys + wg - 1) / wg;
    if (blocks > 0x7fffffffLL) return set_error(GS_EINVAL, "batch too large for one launch");
    if (diag) {
        KernelArgsT<typename Fam::T> b = a;
        b.stage = !holds_state<Fam>::value && a.order == nullptr && wg % 64 == 0 &&
                  reinterpret_cast<uintptr_t>(a.S) % 16 == 0 && reinterpret_cast<uintptr_t>(a.X) % 16 == 0;
        const size_t lds = b.stage ? size_t(wg / 64) * small_stage_bytes<typename Fam::T, Fam::N>() : 0;
        hipLaunchKernelGGL((solve_small_kernel<Fam, GLOB, MIXED, REFINE, true, REUSE>), dim3(unsigned(blocks)), dim3(wg),
                           lds, st, b);
    }
    else
        hipLaunchKernelGGL((solve_small_kernel<Fam, GLOB, MIXED, REFINE, false, REUSE>), dim3(unsigned(blocks)),
                           dim3(wg), 0, st, a);
    return GS_OK;
}

// The launch of one globalization's kernels for family Fam. A Float32 batch has no mixed form: its
// factor is in its working precision already (factor_eltype(LUFactorization{Float32}, Float32)).
template <class Fam, int GLOB, bool REUSE>
int launch_glob(const KernelArgsT<typename Fam::T>& a, bool mixed, bool refine, bool diag, int wg, hipStream_t st) {
    if constexpr (sizeof(typename Fam::T) == 8) {
        if (mixed && !refine) return launch_diag<Fam, GLOB, true, false, REUSE>(a, diag, wg, st);
        if (mixed) return launch_diag<Fam, GLOB, true, true, REUSE>(a, diag, wg, st);
    }
    if (!refine) return launch_diag<Fam, GLOB, false, false, REUSE>(a, diag, wg, st);
    return launch_diag<Fam, GLOB, false, true, REUSE>(a, diag, wg, st);
}

// The family dispatch of one globalization (GS_FAM_* → Fam); EFAMILY for a family without a
// lane-per-system form.
template <int GLOB, bool REUSE, class T>
int launch_small_family(int family, const KernelArgsT<T>& a, bool mixed, bool refine, bool diag, int wg,
                        hipStream_t st) {
    switch (family) {
        case GS_FAM_COUPLED2: return launch_glob<FamCoupled2T<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        case GS_FAM_HETERO2: return launch_glob<FamHetero2T<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        case GS_FAM_SCALAR_SQUARE: return launch_glob<FamScalarSquareT<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        case GS_FAM_TANH_LINEAR3: return launch_glob<FamTanhLinear3T<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        case GS_FAM_QUAD_MAP2: return launch_glob<FamQuadMap2T<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        case GS_FAM_HENON_HEILES_MIDPOINT:
            return launch_glob<FamHenonHeilesMidpointT<T>, GLOB, REUSE>(a, mixed, refine, diag, wg, st);
        default: return set_error(GS_EFAMILY, "problem family is not registered");
    }
}

// Defined in small_reuse.hip: Static and Backtracking with reuse = k.
int launch_small_reuse(int family, int glob, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                       hipStream_t st);

// Defined in small_dogleg.hip: DogLeg.
int launch_small_dogleg(int family, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg, hipStream_t st);

// Defined in small_deflated.hip: Static, Backtracking and DogLeg on a Deflated batch (FamDeflatedT),
// with or without reuse = k (not with DogLeg).
int launch_small_deflated(int family, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag,
                          int wg, hipStream_t st);

// Defined in small_stationarity.hip: Static, Backtracking and DogLeg on a StationarityProblem batch
// of a registered action (FamStationarityT), |E| = 2 or 6, with or without reuse = k (not with DogLeg).
int launch_small_stationarity(int action, int ne, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                              bool diag, int wg, hipStream_t st);

// Defined in small_reparam.hip: Static, Backtracking and DogLeg on a Reparametrized batch
// (FamReparamT) with a linear chart (families 1–5) or the polar chart (the n = 2 families), with or
// without reuse = k (not with DogLeg).
int launch_small_reparam(int family, int chart, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine,
                         bool diag, int wg, hipStream_t st);

// Defined in small_stationarity_f32.hip and small_reparam_f32.hip: the Float32 forms of the two
// batches above (no mixed form; the Reparametrized batch with the linear chart only).
int launch_small_stationarity_f32(int action, int ne, int glob, bool reuse, const KernelArgsT<float>& a, bool refine,
                                  bool diag, int wg, hipStream_t st);
int launch_small_reparam_f32(int family, int chart, int glob, bool reuse, const KernelArgsT<float>& a, bool refine,
                             bool diag, int wg, hipStream_t st);

// Defined in small_dense.hip: the dense family at n = 2, 4, 6, 8 (FamDenseSinSmallT, A in
// a.fam.data), in FP64 and Float32; glob is kStatic, kBacktracking or kDogLeg.
int launch_small_dense(int n, int glob, bool reuse, const KernelArgs& a, bool mixed, bool refine, bool diag, int wg,
                       hipStream_t st);
int launch_small_dense_f32(int n, int glob, bool reuse, const KernelArgsT<float>& a, bool refine, bool diag, int wg,
                           hipStream_t st);

// Defined in small_f32.hip: every form of a Float32 batch.
int launch_small_f32(int family, int glob, bool reuse, const KernelArgsT<float>& a, bool refine, bool diag, int wg,
                     hipStream_t st);

}  // namespace gs
