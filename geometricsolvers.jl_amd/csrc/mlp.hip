// libgsolve_hip, learned start: a multilayer perceptron that predicts the Newton start of the
// implicit-midpoint step of Hénon–Heiles, with its inference and training kernels on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation).
//
// The network g: R^32 → R^16 (5 inputs and 4 outputs used) has three hidden layers of 128 ReLU
// units. It predicts the correction to the explicit midpoint step x̃ = x + h f(x + (h/2) f(x)),
// which differs from the implicit midpoint step by O(h³):
//     x̂₊ = x̃ + h³ (o_j σ_j + μ_j),   o = g(u),  u_j = (x_j − s_j) c_j (j < 4), u_4 = (h − s_4) c_4,
// so that the network's error (bounded below by the bfloat16 rounding of its inputs) enters the
// start with a factor h³. Newton in FP64 then polishes x̂₊; the network has no counterpart in the
// reference.
//
// Parameters live in one FP32 array (W1, b1, W2, b2, W3, b3, W4, b4; W_l row-major in × out), and
// the inference kernel reads a BF16 copy of the weights packed as the MFMA A operands of the
// transposed products Wᵀ aᵀ, which the Adam kernel writes after each update.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/gsolve.h"
#include "families.hpp"
#include "mid_solver.hpp"

namespace gs {

int set_error(int code, const std::string& msg);
// One call at a time per device, held through the synchronisation (defined with gs_solve_batch_f64).
std::unique_lock<std::recursive_mutex> device_call_lock();

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = GS_MLP_IN, kHid = GS_MLP_HIDDEN, kOut = GS_MLP_OUT;

// Offsets of the parameter array.
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kHid * kHid, kW4 = kB3 + kHid, kB4 = kW4 + kHid * kOut, kNParam = kB4 + kOut;
static_assert(kNParam == GS_MLP_NPARAM, "parameter count");
// Offsets of the packed BF16 weights (the biases are not packed).
constexpr int kP1 = 0, kP2 = kP1 + kIn * kHid, kP3 = kP2 + kHid * kHid, kP4 = kP3 + kHid * kHid,
              kNPacked = kP4 + kHid * kOut;
static_assert(kNPacked == GS_MLP_NPACKED, "packed weight count");

// The MFMA operand layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, g = l / 16):
//   A (16 × 32): lane l holds row l % 16, columns 8 g … 8 g + 7;
//   B (32 × 16): lane l holds column l % 16, rows 8 g … 8 g + 7;
//   D (16 × 16): lane l holds column l % 16, rows 4 g … 4 g + 3.
// Inference computes each layer transposed, zᵀ = Wᵀ aᵀ over 16 systems at a time: A is a 16-unit
// block of Wᵀ, B the previous activations (columns = systems) and D the new ones. D's rows 4 g … 4 g + 3
// of unit blocks 2 s and 2 s + 1 are then exactly the eight B rows of lane group g in k-step s, once the
// k index of every k-step is permuted: slot j of lane group g stands for unit
//     32 s + 4 g + j (j < 4),   32 s + 16 + 4 g + (j − 4) (j ≥ 4),
// so that activations pass from layer to layer in registers. The packed weights are the A operands
// under that permutation, in the order (k-step, unit block, lane, slot).
GS_DEV int packed_index(int k, int col, int out) {
    const int ks = k >> 5, kk = k & 31, ob = col >> 4;
    const int g = (kk & 15) >> 2, slot = (kk & 3) | ((kk & 16) >> 2);
    const int lane = (col & 15) + 16 * g;
    return ((ks * (out >> 4) + ob) * 64 + lane) * 8 + slot;
}

GS_DEV float relu(float z) { return z > 0.0f ? z : 0.0f; }

// ReLU of two FP32 values into a packed BF16 pair: round each to BF16, then clear every value whose
// sign bit is set, as an int16 maximum with 0 (one packed integer instruction for the pair).
GS_DEV uint32_t relu_bf16x2(float a, float b) {
    const bf16x2 p = __builtin_convertvector(f32x2{a, b}, bf16x2);
    const s16x2 v = __builtin_elementwise_max(__builtin_bit_cast(s16x2, p), s16x2{0, 0});
    return __builtin_bit_cast(uint32_t, v);
}

// Hénon–Heiles vector field f(q₁, q₂, p₁, p₂) = (p₁, p₂, −q₁ − 2 q₁ q₂, −q₂ − q₁² + q₂²).
GS_DEV void hh_field(const double (&x)[4], double (&f)[4]) {
    f[0] = x[2];
    f[1] = x[3];
    f[2] = -x[0] - 2.0 * x[0] * x[1];
    f[3] = (-x[1] - x[0] * x[0]) + x[1] * x[1];
}

// The explicit midpoint step x̃ = x + h f(x + (h/2) f(x)), the base of the prediction.
GS_DEV void hh_rk2(const double (&x)[4], double h, double (&xt)[4]) {
    double f[4], m[4], fm[4];
    hh_field(x, f);
#pragma unroll
    for (int j = 0; j < 4; ++j) m[j] = x[j] + (0.5 * h) * f[j];
    hh_field(m, fm);
#pragma unroll
    for (int j = 0; j < 4; ++j) xt[j] = x[j] + h * fm[j];
}

struct MlpNorm {
    float in_shift[8], in_scale[8], out_shift[4], out_scale[4];
};

MlpNorm norm_of(const gs_mlp& n) {
    MlpNorm m;
    for (int j = 0; j < 8; ++j) {
        m.in_shift[j] = n.in_shift[j];
        m.in_scale[j] = n.in_scale[j];
    }
    for (int j = 0; j < 4; ++j) {
        m.out_shift[j] = n.out_shift[j];
        m.out_scale[j] = n.out_scale[j];
    }
    return m;
}

// ---- inference: the start of every system ----------------------------------------------------------

struct StartArgs {
    const double* Xk;
    double* X0;
    int64_t nsys;
    double h;
    const __bf16* packed;
    const float* params;
    MlpNorm nm;
};

constexpr int kStartWaves = 8;
constexpr int kTiles = 4;  // 16-system tiles per wave pass: each weight fragment read feeds kTiles MFMAs
constexpr int kChunk = 16 * kTiles;
static_assert(kChunk == 64, "one lane per system of a chunk");
// LDS of a workgroup: packed weights, biases, then per wave the corrections of its kChunk systems.
constexpr size_t kLdsWeights = size_t(kNPacked) * 2;
constexpr size_t kLdsBias = size_t(3 * kHid + kOut) * 4;
constexpr size_t kLdsCorr = size_t(kChunk) * 4 * 4;
constexpr size_t kStartLds = kLdsWeights + kLdsBias + size_t(kStartWaves) * kLdsCorr;

// The workgroup's view of its LDS.
struct MlpLds {
    const __bf16* w;
    const float* bias;
    float* corr;   // this wave's kChunk × 4 outputs
};

// Load the packed weights and biases into LDS (the whole workgroup) and return this wave's view.
GS_DEV MlpLds mlp_lds_setup(unsigned char* lds, const __bf16* packed, const float* params) {
    __bf16* w = reinterpret_cast<__bf16*>(lds);
    float* bias = reinterpret_cast<float*>(lds + kLdsWeights);
    const uint4* src = reinterpret_cast<const uint4*>(packed);
    uint4* dst = reinterpret_cast<uint4*>(w);
    for (int i = threadIdx.x; i < kNPacked / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        bias[i] = params[kB1 + i];
        bias[kHid + i] = params[kB2 + i];
        bias[2 * kHid + i] = params[kB3 + i];
    }
    for (int i = threadIdx.x; i < kOut; i += blockDim.x) bias[3 * kHid + i] = params[kB4 + i];
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    MlpLds v;
    v.w = w;
    v.bias = bias;
    v.corr = reinterpret_cast<float*>(lds + kLdsWeights + kLdsBias + size_t(wave) * kLdsCorr);
    return v;
}

GS_DEV void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

GS_DEV double lane_value(double v, int src) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __shfl(int(uint32_t(u)), src), hi = __shfl(int(uint32_t(u >> 32)), src);
    return __builtin_bit_cast(double, (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// The network on TILES × 16 systems of a chunk, from system `base` on. Lane l passes the state x of
// system l of the chunk (any finite value past the batch); the outputs o + b4 of system b land in
// L.corr[4 b + j]. Tile t holds systems base + 16 t … base + 16 t + 15 as the B columns; every weight
// fragment is read from LDS once and multiplies all TILES tiles (the arithmetic of a system does not
// depend on TILES).
template <int TILES> GS_DEV void mlp_tiles(const MlpLds& L, const MlpNorm& nm, double h, const double (&x)[4], int base) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    constexpr int kKs = kHid / 32, kOb = kHid / 16;
    // activations as packed BF16 pairs: nxt[t][s] is the B operand of tile t in k-step s
    u32x4 act[TILES][kKs], nxt[TILES][kKs];
    const float uh = (float(h) - nm.in_shift[4]) * nm.in_scale[4];
    // layer 1 inputs: slots 0..3 of lane group 0 are the features of x, slot 0 of group 1 is h's
    bf16x8 in[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t) {
        double xt[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) xt[j] = lane_value(x[j], base + 16 * t + r);
#pragma unroll
        for (int j = 0; j < 8; ++j) in[t][j] = __bf16(0.0f);
        if (g == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) in[t][j] = __bf16((float(xt[j]) - nm.in_shift[j]) * nm.in_scale[j]);
        } else if (g == 1) {
            in[t][0] = __bf16(uh);
        }
    }
    // The bias of unit block ob, rows 4 g … 4 g + 3: the C operand of the block's first MFMA.
    auto bias_of = [&](const float* b, int ob) { return *reinterpret_cast<const f32x4*>(b + ob * 16 + 4 * g); };
    // ReLU of unit block ob into the B operands of the next layer (blocks 2 s, 2 s + 1 → k-step s)
    auto activate = [&](const f32x4 (&acc)[TILES], int ob) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            nxt[t][ob >> 1][(ob & 1) * 2] = relu_bf16x2(acc[t][0], acc[t][1]);
            nxt[t][ob >> 1][(ob & 1) * 2 + 1] = relu_bf16x2(acc[t][2], acc[t][3]);
        }
    };
#pragma unroll
    for (int ob = 0; ob < kOb; ++ob) {
        const bf16x8 wv = *reinterpret_cast<const bf16x8*>(L.w + kP1 + (ob * 64 + lane) * 8);
        const f32x4 bv = bias_of(L.bias, ob);
        f32x4 acc[TILES];
#pragma unroll
        for (int t = 0; t < TILES; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, in[t], bv, 0, 0, 0);
        activate(acc, ob);
    }
    // a hidden layer, one unit block at a time: only that block's accumulators are live
    auto hidden = [&](int poff, const float* b) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
#pragma unroll
            for (int ks = 0; ks < kKs; ++ks) act[t][ks] = nxt[t][ks];
        }
#pragma unroll
        for (int ob = 0; ob < kOb; ++ob) {
            const f32x4 bv = bias_of(b, ob);
            f32x4 acc[TILES];
#pragma unroll
            for (int ks = 0; ks < kKs; ++ks) {
                const bf16x8 wv = *reinterpret_cast<const bf16x8*>(L.w + poff + ((ks * kOb + ob) * 64 + lane) * 8);
#pragma unroll
                for (int t = 0; t < TILES; ++t)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, __builtin_bit_cast(bf16x8, act[t][ks]),
                                                                      ks == 0 ? bv : acc[t], 0, 0, 0);
            }
            activate(acc, ob);
        }
    };
    hidden(kP2, L.bias + kHid);
    hidden(kP3, L.bias + 2 * kHid);
    f32x4 o[TILES];
    const f32x4 b4 = bias_of(L.bias + 3 * kHid, 0);
#pragma unroll
    for (int ks = 0; ks < kKs; ++ks) {
        const bf16x8 wv = *reinterpret_cast<const bf16x8*>(L.w + kP4 + (ks * 64 + lane) * 8);
#pragma unroll
        for (int t = 0; t < TILES; ++t)
            o[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, __builtin_bit_cast(bf16x8, nxt[t][ks]), ks == 0 ? b4 : o[t],
                                                           0, 0, 0);
    }
    // outputs 0..3 are rows 0..3 of D: lane group 0
    if (g == 0) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) *reinterpret_cast<f32x4*>(L.corr + (base + 16 * t + r) * 4) = o[t];
    }
    wave_lds_fence();
}

// The network on the kChunk systems of one wave pass (kTiles tiles).
GS_DEV void mlp_chunk(const MlpLds& L, const MlpNorm& nm, double h, const double (&x)[4]) {
    mlp_tiles<kTiles>(L, nm, h, x, 0);
}

GS_DEV void load_state(const double* X, int64_t s, double (&x)[4]) {
    const double2* xp = reinterpret_cast<const double2*>(X + 4 * s);
    const double2 u = xp[0], v = xp[1];
    x[0] = u.x;
    x[1] = u.y;
    x[2] = v.x;
    x[3] = v.y;
}

// x̂ = x̃ + h³ (o σ + μ) from a row of outputs.
GS_DEV void predicted_start(const double (&x)[4], double h, const float* o, const MlpNorm& nm, double (&y)[4]) {
    double xt[4];
    hh_rk2(x, h, xt);
    const double h3 = (h * h) * h;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double corr = double(o[j] * nm.out_scale[j] + nm.out_shift[j]);
        y[j] = xt[j] + h3 * corr;
    }
}

// One workgroup of 8 waves per CU holds the packed weights and biases in LDS; each wave runs chunks
// of kChunk systems through the four layers, its activations in registers.
__global__ __launch_bounds__(64 * kStartWaves) void mlp_start_kernel(const StartArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const MlpLds L = mlp_lds_setup(lds, a.packed, a.params);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nchunk = (a.nsys + kChunk - 1) / kChunk;
    for (int64_t c = int64_t(blockIdx.x) * kStartWaves + wave; c < nchunk; c += int64_t(gridDim.x) * kStartWaves) {
        const int64_t s = c * kChunk + lane;
        const bool live = s < a.nsys;
        double x[4] = {0.0, 0.0, 0.0, 0.0};
        if (live) load_state(a.Xk, s, x);
        mlp_chunk(L, a.nm, a.h, x);
        if (live) {
            double y[4];
            predicted_start(x, a.h, L.corr + lane * 4, a.nm, y);
            double2* yp = reinterpret_cast<double2*>(a.X0 + 4 * s);
            yp[0] = double2{y[0], y[1]};
            yp[1] = double2{y[2], y[3]};
        }
        wave_lds_fence();
    }
}

// ---- the fused solve: learned start, Newton from it, plain Newton where that fails ---------------

struct LearnedArgs {
    StartArgs st;
    DevParams q;    // the method of the solve from the learned start
    DevParams qf;   // the fallback: plain Newton from x_k
    gs_status_f64* S;
    uint32_t* counters;  // [0] systems that fell back, [1] waves whose hand-over wait ran out
};

GS_DEV void store_status_learned(gs_status_f64* dst, uint8_t code, int32_t iterations, double fnorm, double stepnorm,
                                 int32_t failures, bool promoted, double gradratio, double rate, bool contracted) {
    uint64_t* w = reinterpret_cast<uint64_t*>(dst);
    w[0] = uint64_t(code) | (uint64_t(uint32_t(iterations)) << 32);
    w[1] = __builtin_bit_cast(uint64_t, fnorm);
    w[2] = __builtin_bit_cast(uint64_t, stepnorm);
    w[3] = uint64_t(uint32_t(failures)) | (uint64_t(promoted ? 1 : 0) << 32);
    w[4] = uint64_t(uint32_t(-1));
    w[5] = __builtin_bit_cast(uint64_t, dnan());
    w[6] = __builtin_bit_cast(uint64_t, gradratio);
    w[7] = __builtin_bit_cast(uint64_t, rate);
    w[8] = uint64_t(contracted ? 1 : 0);
}

// The workgroup of the fused solve: four producer waves (one per SIMD) run the network on BF16 MFMA
// and eight consumer waves (two per SIMD) run Newton in FP64 on the VALU, so that the two pipes work
// at once. A workgroup takes groups of kWsCons chunks (64 systems each) in rounds; in round r,
// producer p forms the predicted starts of chunks 2p and 2p + 1 of the round's group into LDS slots
// (round parity × kWsCons + chunk), and consumer c solves chunk c from its slot. Slot flags: full =
// the rounds produced into the slot, done = the rounds its consumer has read; a producer refills a
// slot once the round two back has been read.
constexpr int kWsProd = 4, kWsCons = 8, kWsWaves = kWsProd + kWsCons;
constexpr int kWsSlots = 2 * kWsCons;
constexpr size_t kWsSlotBytes = size_t(kChunk) * 4 * sizeof(double);
constexpr size_t kWsCorr = kLdsWeights + kLdsBias;
constexpr size_t kWsSlotOff = kWsCorr + size_t(kWsProd) * kLdsCorr;
constexpr size_t kWsFlagOff = kWsSlotOff + size_t(kWsSlots) * kWsSlotBytes;
constexpr size_t kWsLds = kWsFlagOff + size_t(2 * kWsSlots) * sizeof(int);
static_assert(kWsLds <= 160 * 1024, "the fused solve's LDS");
static_assert(kWsProd * 2 == kWsCons, "two chunks per producer and round");
// A wait gives up after this many polls (each an LDS read and a short sleep: over a second), so that
// a broken hand-over ends the kernel with an error instead of hanging it.
constexpr int kWsSpinLimit = 1 << 24;

// Wait until *flag ≥ v (acquire at workgroup scope); false when the wait runs out.
GS_DEV bool wait_flag(const int* flag, int v) {
    for (int n = 0; n < kWsSpinLimit; ++n) {
        if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= v) return true;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// Publish v in *flag after the wave's earlier LDS accesses (release at workgroup scope).
GS_DEV void post_flag(int* flag, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Producer p: the network and the predicted starts of chunks 2p and 2p + 1 of every round.
GS_DEV void learned_produce(const LearnedArgs& a, const MlpLds& L, int p, int64_t nround, int64_t nchunk, double* slots,
                            const int* done, int* full) {
    const int lane = threadIdx.x & 63;
    for (int64_t r = 0; r < nround; ++r) {
        const int64_t g = int64_t(blockIdx.x) + r * gridDim.x;
#pragma unroll 1
        for (int i = 0; i < 2; ++i) {
            const int c = 2 * p + i;
            const int64_t chunk = g * kWsCons + c;
            const int64_t s = chunk * kChunk + lane;
            double y[4] = {0.0, 0.0, 0.0, 0.0};
            if (chunk < nchunk) {  // uniform over the wave: the MFMAs run with every lane
                {
                    double xk[4] = {0.0, 0.0, 0.0, 0.0};
                    if (s < a.st.nsys) load_state(a.st.Xk, s, xk);
                    // two tiles at a time: half the activations' registers of mlp_chunk
                    mlp_tiles<2>(L, a.st.nm, a.st.h, xk, 0);
                    mlp_tiles<2>(L, a.st.nm, a.st.h, xk, 32);
                }
                if (s < a.st.nsys) {
                    double xk[4];  // read again (from L2) rather than held through the network
                    load_state(a.st.Xk, s, xk);
                    predicted_start(xk, a.st.h, L.corr + lane * 4, a.st.nm, y);
                }
                wave_lds_fence();  // the outputs are read before the next chunk's network overwrites them
            }
            const int slot = int(r & 1) * kWsCons + c;
            if (!wait_flag(done + slot, int(r) - 1)) {
                if (lane == 0) atomicAdd(a.counters + 1, 1u);
                return;
            }
            double2* dst = reinterpret_cast<double2*>(slots + (size_t(slot) * kChunk + lane) * 4);
            dst[0] = double2{y[0], y[1]};
            dst[1] = double2{y[2], y[3]};
            post_flag(full + slot, int(r) + 1);
        }
    }
}

// Consumer c: chunk c of every round, from the learned start; a system whose solve does not end in
// SUCCESS solves again from x_k with plain Newton, and that status is the system's.
template <bool MIXED, bool REFINE>
GS_DEV void learned_consume(const LearnedArgs& a, int c, int64_t nround, const double* slots, const int* full,
                            int* done) {
    const int lane = threadIdx.x & 63;
    FamArgs fa;
    fa.P = reinterpret_cast<const char*>(a.st.Xk);
    fa.p_stride = 32;
    fa.shared[0] = a.st.h;
    for (int k = 1; k < 8; ++k) fa.shared[k] = 0.0;
    for (int64_t r = 0; r < nround; ++r) {
        const int64_t g = int64_t(blockIdx.x) + r * gridDim.x;
        const int64_t s = (g * kWsCons + c) * kChunk + lane;
        const int slot = int(r & 1) * kWsCons + c;
        if (!wait_flag(full + slot, int(r) + 1)) {
            if (lane == 0) atomicAdd(a.counters + 1, 1u);
            return;
        }
        const double2* src = reinterpret_cast<const double2*>(slots + (size_t(slot) * kChunk + lane) * 4);
        const double2 u = src[0], v = src[1];
        post_flag(done + slot, int(r) + 1);
        if (s < a.st.nsys) {
            double x[4] = {u.x, u.y, v.x, v.y};
            FamHenonHeilesMidpoint fam;
            fam.load(fa, s);
            SmallSolve<FamHenonHeilesMidpoint, kStatic, MIXED, REFINE, true> sol(fam, a.q);
            const auto R = sol.solve(x);
            double2* xp = reinterpret_cast<double2*>(a.st.X0 + 4 * s);
            if (R.code == kSuccess) {
                store_status_learned(a.S + s, R.code, R.iterations, R.fnorm, R.stepnorm, R.failures, R.promoted,
                                     R.gradratio, R.rate, R.contracted);
            } else {
                load_state(a.st.Xk, s, x);
                SmallSolve<FamHenonHeilesMidpoint, kStatic, false, false, true> fb(fam, a.qf);
                const auto Rf = fb.solve(x);
                atomicAdd(a.counters, 1u);
                store_status_learned(a.S + s, Rf.code, Rf.iterations, Rf.fnorm, Rf.stepnorm, Rf.failures, Rf.promoted,
                                     Rf.gradratio, Rf.rate, Rf.contracted);
            }
            xp[0] = double2{x[0], x[1]};
            xp[1] = double2{x[2], x[3]};
        }
    }
}

template <bool MIXED, bool REFINE>
__global__ __launch_bounds__(64 * kWsWaves) void learned_solve_kernel(const LearnedArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    int* full = reinterpret_cast<int*>(lds + kWsFlagOff);
    int* done = full + kWsSlots;
    if (threadIdx.x < 2 * kWsSlots) full[threadIdx.x] = 0;  // full and done (mlp_lds_setup synchronises)
    MlpLds L = mlp_lds_setup(lds, a.st.packed, a.st.params);
    const int wave = threadIdx.x >> 6;
    double* slots = reinterpret_cast<double*>(lds + kWsSlotOff);
    const int64_t nchunk = (a.st.nsys + kChunk - 1) / kChunk;
    const int64_t ngroup = (nchunk + kWsCons - 1) / kWsCons;
    const int64_t nround = int64_t(blockIdx.x) < ngroup ? (ngroup - 1 - blockIdx.x) / gridDim.x + 1 : 0;
    if (wave < kWsProd) {
        L.corr = reinterpret_cast<float*>(lds + kWsCorr + size_t(wave) * kLdsCorr);
        learned_produce(a, L, wave, nround, nchunk, slots, done, full);
    } else {
        learned_consume<MIXED, REFINE>(a, wave - kWsProd, nround, slots, full, done);
    }
}

// ---- training ----------------------------------------------------------------------------------------

// C = op(A) · op(B), one wave per 16 × 16 block of C (M and N multiples of 16), K in steps of 32; A(i, k) = A[i as0 + k as1],
// B(k, j) = B[k bs0 + j bs1], converted to BF16 on load (FP32 master weights or BF16 activations).
// Epilogues: kPlain C = acc (FP32), kReluAct Z = acc + b → act = relu(Z) (BF16), kBias C = acc + b.
enum { kPlain = 0, kReluAct = 1, kBias = 2 };

template <class TA, class TB> struct GemmArgs {
    const TA* A;
    int64_t as0, as1;
    const TB* B;
    int64_t bs0, bs1;
    int M, N, K;
    int mode;
    const float* bias;
    float* C;
    __bf16* Cb;
    int ldc;
};

template <class T> GS_DEV __bf16 to_bf16(T v) { return __bf16(float(v)); }
template <> GS_DEV __bf16 to_bf16<__bf16>(__bf16 v) { return v; }

template <class TA, class TB> __global__ __launch_bounds__(256) void gemm_kernel(const GemmArgs<TA, TB> g) {
    const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int nbn = g.N / 16;
    if (wave >= int64_t(g.M / 16) * nbn) return;
    const int bi = int(wave / nbn), bj = int(wave % nbn);
    const int r = lane & 15, kg = lane >> 4;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const TA* ap = g.A + int64_t(bi * 16 + r) * g.as0;
    const TB* bp = g.B + int64_t(bj * 16 + r) * g.bs1;
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            // K need not be a multiple of 32: the entries past K are zeros, not the next row's
            const int64_t k = k0 + 8 * kg + j;
            const bool in = k < g.K;
            av[j] = in ? to_bf16(ap[in ? k * g.as1 : 0]) : __bf16(0.0f);
            bv[j] = in ? to_bf16(bp[in ? k * g.bs0 : 0]) : __bf16(0.0f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    const int col = bj * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = bi * 16 + 4 * kg + i;
        float v = acc[i];
        if (g.mode != kPlain) v += g.bias[col];
        if (g.mode == kReluAct) {
            g.C[int64_t(row) * g.ldc + col] = v;
            g.Cb[int64_t(row) * g.ldc + col] = __bf16(relu(v));
        } else {
            g.C[int64_t(row) * g.ldc + col] = v;
        }
    }
}

template <class TA, class TB> void gemm(const GemmArgs<TA, TB>& g, hipStream_t st) {
    const int64_t waves = int64_t(g.M / 16) * (g.N / 16);
    hipLaunchKernelGGL((gemm_kernel<TA, TB>), dim3(unsigned((waves + 3) / 4)), dim3(256), 0, st, g);
}

// The features and normalised targets of a minibatch: row b of U (BF16, kIn wide) and T (FP32, kOut
// wide) from system idx[b] of (Xk, X1).
__global__ void features_kernel(const double* Xk, const double* X1, const int64_t* idx, int batch, double h,
                                MlpNorm nm, __bf16* U, float* T) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const int64_t s = idx[b];
    double x[4], y[4], xt[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        x[j] = Xk[4 * s + j];
        y[j] = X1[4 * s + j];
    }
    hh_rk2(x, h, xt);
    const double h3 = (h * h) * h;
    for (int j = 0; j < kIn; ++j) {
        float u = 0.0f;
        if (j < 4) u = (float(x[j]) - nm.in_shift[j]) * nm.in_scale[j];
        if (j == 4) u = (float(h) - nm.in_shift[4]) * nm.in_scale[4];
        U[int64_t(b) * kIn + j] = __bf16(u);
    }
    for (int j = 0; j < kOut; ++j) {
        float t = 0.0f;
        if (j < 4) {
            const double corr = (y[j] - xt[j]) / h3;
            t = (float(corr) - nm.out_shift[j]) / nm.out_scale[j];
        }
        T[int64_t(b) * kOut + j] = t;
    }
}

// Loss ½·mean over the batch and the 4 outputs of (Z4 − T)² (accumulated into *loss) and its
// gradient dZ4 = (Z4 − T) / (4 · batch) (zero on the padding columns).
__global__ void loss_kernel(const float* Z4, const float* T, int batch, float* dZ4, float* loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < batch * kOut) {
        const int j = i % kOut;
        const float d = j < 4 ? Z4[i] - T[i] : 0.0f;
        dZ4[i] = d / float(4 * batch);
        l = 0.5f * d * d / float(4 * batch);
    }
    // wave sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// dZ = dA ⊙ [Z > 0] for a ReLU layer (dA FP32 from the GEMM, Z the layer's FP32 pre-activation).
__global__ void relu_backward_kernel(const float* dA, const float* Z, int64_t n, float* dZ) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = Z[i] > 0.0f ? dA[i] : 0.0f;
}

// db_j = Σ_b dZ[b, j] (one workgroup per column).
__global__ __launch_bounds__(256) void colsum_kernel(const float* dZ, int batch, int width, float* db) {
    __shared__ float part[256];
    const int j = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < batch; b += blockDim.x) s += dZ[int64_t(b) * width + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = part[0];
}

__global__ void fill_kernel(float* p, int64_t n, float v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The fused solve's two counters (fallbacks, hand-over time-outs) copied into pinned host memory by
// the kernel queued right behind the solve, so the host reads them after the one synchronisation
// without a copy of its own, and cleared for the next solve.
__global__ void publish_counters_kernel(uint32_t* counters, uint32_t* host) {
    if (threadIdx.x < 2) {
        host[threadIdx.x] = counters[threadIdx.x];
        counters[threadIdx.x] = 0;
    }
    __threadfence_system();
}

// Adam (Kingma & Ba) on the FP32 parameters, then the packed BF16 copy of the weights.
__global__ void adam_kernel(float* p, const float* grad, float* m, float* v, int64_t t, float lr, float b1, float b2,
                            float eps, __bf16* packed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNParam) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - __builtin_powf(b1, float(t)));
    const float vh = vi / (1.0f - __builtin_powf(b2, float(t)));
    const float pi = p[i] - lr * mh / (__builtin_sqrtf(vh) + eps);
    p[i] = pi;
    // the weights' packed BF16 copy
    int off = -1, k = 0, col = 0, out = kHid;
    if (i < kB1) { off = kP1; k = i / kHid; col = i % kHid; }
    else if (i >= kW2 && i < kB2) { off = kP2; k = (i - kW2) / kHid; col = (i - kW2) % kHid; }
    else if (i >= kW3 && i < kB3) { off = kP3; k = (i - kW3) / kHid; col = (i - kW3) % kHid; }
    else if (i >= kW4 && i < kB4) { off = kP4; k = (i - kW4) / kOut; col = (i - kW4) % kOut; out = kOut; }
    if (off >= 0) packed[off + packed_index(k, col, out)] = __bf16(pi);
}

// Repack all weights (after the caller sets the parameters directly).
__global__ void pack_kernel(const float* p, __bf16* packed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNParam) return;
    int off = -1, k = 0, col = 0, out = kHid;
    if (i < kB1) { off = kP1; k = i / kHid; col = i % kHid; }
    else if (i >= kW2 && i < kB2) { off = kP2; k = (i - kW2) / kHid; col = (i - kW2) % kHid; }
    else if (i >= kW3 && i < kB3) { off = kP3; k = (i - kW3) / kHid; col = (i - kW3) % kHid; }
    else if (i >= kW4 && i < kB4) { off = kP4; k = (i - kW4) / kOut; col = (i - kW4) % kOut; out = kOut; }
    if (off >= 0) packed[off + packed_index(k, col, out)] = __bf16(p[i]);
}

// Training workspace (per device), sized for the largest batch seen.
struct MlpWorkspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool clear = false;  // the fused solve's counters hold zeros (publish_counters_kernel cleared them)
};
std::mutex g_mlp_mutex;
MlpWorkspace g_mlp_ws[64];

MlpWorkspace g_mlp_counter[64];  // the fused solve's fallback counter (its own, so training can overlap)

// The pinned, device-mapped host words publish_counters_kernel writes (per device).
struct MlpPublish {
    uint32_t* host = nullptr;
    uint32_t* dev = nullptr;
};
MlpPublish g_mlp_pub[64];

int mlp_publish(MlpPublish* out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_mlp_mutex);
    MlpPublish& w = g_mlp_pub[dev];
    if (!w.host) {
        void* h = nullptr;
        if (hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return set_error(GS_EHIP, "hipHostMalloc of the counter read-back failed");
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
            (void)hipHostFree(h);
            return set_error(GS_EHIP, "hipHostGetDevicePointer of the counter read-back failed");
        }
        w.host = static_cast<uint32_t*>(h);
        w.dev = static_cast<uint32_t*>(d);
    }
    *out = w;
    return GS_OK;
}

int mlp_workspace(size_t bytes, void** out, bool counter = false) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_mlp_mutex);
    MlpWorkspace& w = counter ? g_mlp_counter[dev] : g_mlp_ws[dev];
    if (w.bytes < bytes) {
        if (w.ptr) (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
        if (hipMalloc(&w.ptr, bytes) != hipSuccess) return set_error(GS_EHIP, "hipMalloc of the training workspace failed");
        w.bytes = bytes;
        w.clear = false;
    }
    *out = w.ptr;
    return GS_OK;
}

int check(hipError_t e, const char* what) {
    if (e != hipSuccess) return set_error(GS_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return GS_OK;
}

}  // namespace

// Free the training workspace, the fallback counter and its pinned read-back of every device (gs_shutdown; no call may be
// in flight).
void release_mlp() {
    std::lock_guard<std::mutex> lock(g_mlp_mutex);
    for (int d = 0; d < 64; ++d) {
        for (MlpWorkspace* w : {&g_mlp_ws[d], &g_mlp_counter[d]}) {
            if (!w->ptr) continue;
            if (hipSetDevice(d) == hipSuccess) (void)hipFree(w->ptr);
            *w = MlpWorkspace{};
        }
        if (g_mlp_pub[d].host) {
            if (hipSetDevice(d) == hipSuccess) (void)hipHostFree(g_mlp_pub[d].host);
            g_mlp_pub[d] = MlpPublish{};
        }
    }
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_mlp_start_henon_heiles(const gs_mlp* net, int64_t nsys, const double* Xk, double h, double* X0, void* stream) {
    if (!net || !net->params || !net->packed) return set_error(GS_EINVAL, "the network must not be NULL");
    if (nsys < 0) return set_error(GS_EINVAL, "nsys must be ≥ 0");
    if (nsys == 0) return GS_OK;
    if (!Xk || !X0) return set_error(GS_EINVAL, "Xk and X0 must not be NULL");
    if ((reinterpret_cast<uintptr_t>(Xk) % 16) || (reinterpret_cast<uintptr_t>(X0) % 16) ||
        (reinterpret_cast<uintptr_t>(net->packed) % 16))
        return set_error(GS_EINVAL, "Xk, X0 and the packed weights must be 16-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    static bool attr_set[64] = {};
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    if (!attr_set[dev]) {
        int rc = check(hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_start_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(kStartLds)),
                       "hipFuncSetAttribute");
        if (rc) return rc;
        attr_set[dev] = true;
    }
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    StartArgs a;
    a.Xk = Xk;
    a.X0 = X0;
    a.nsys = nsys;
    a.h = h;
    a.packed = reinterpret_cast<const __bf16*>(net->packed);
    a.params = net->params;
    a.nm = norm_of(*net);
    const int64_t nchunk = (nsys + kChunk - 1) / kChunk;
    int64_t blocks = (nchunk + kStartWaves - 1) / kStartWaves;
    if (blocks > cus) blocks = cus;
    hipLaunchKernelGGL(mlp_start_kernel, dim3(unsigned(blocks)), dim3(64 * kStartWaves), kStartLds, st, a);
    return check(hipGetLastError(), "learned-start launch failed");
}

int gs_solve_learned_henon_heiles_f64(const gs_mlp* net, const gs_method* method, int64_t nsys, const double* Xk,
                                      double h, double* X, gs_status_f64* S, const gs_options_f64* options,
                                      int64_t* fallbacks, void* stream) {
    if (!net || !net->params || !net->packed || !method || !options)
        return set_error(GS_EINVAL, "net, method and options must not be NULL");
    if (nsys < 0) return set_error(GS_EINVAL, "nsys must be ≥ 0");
    if (method->linsolve != GS_LU || method->globalization != GS_STATIC || method->reuse != 0)
        return set_error(GS_EFAMILY, "the fused learned solve takes Newton with LU, the Static globalization and no reuse");
    if (method->factor_precision != GS_FACTOR_WORKING && method->factor_precision != GS_FACTOR_F32)
        return set_error(GS_EINVAL, "factor_precision must be GS_FACTOR_WORKING or GS_FACTOR_F32");
    if (fallbacks) *fallbacks = 0;
    if (nsys == 0) return GS_OK;
    if (!Xk || !X || !S) return set_error(GS_EINVAL, "Xk, X and S must not be NULL");
    if ((reinterpret_cast<uintptr_t>(Xk) % 16) || (reinterpret_cast<uintptr_t>(X) % 16) ||
        (reinterpret_cast<uintptr_t>(S) % 8) || (reinterpret_cast<uintptr_t>(net->packed) % 16))
        return set_error(GS_EINVAL, "Xk, X and the packed weights must be 16-byte aligned, S 8-byte aligned");
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto lock = device_call_lock();
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    const int v = method->factor_precision == GS_FACTOR_F32 ? 1 : 0;
    const void* kern = v ? (method->refinement == GS_REFINEMENT ? reinterpret_cast<const void*>(&learned_solve_kernel<true, true>)
                                                                 : reinterpret_cast<const void*>(&learned_solve_kernel<true, false>))
                         : (method->refinement == GS_REFINEMENT ? reinterpret_cast<const void*>(&learned_solve_kernel<false, true>)
                                                                 : reinterpret_cast<const void*>(&learned_solve_kernel<false, false>));
    // one attribute per kernel variant
    static bool attr_kern[64][4] = {};
    const int ki = v * 2 + (method->refinement == GS_REFINEMENT ? 1 : 0);
    if (!attr_kern[dev][ki]) {
        int rc = check(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(kWsLds)),
                       "hipFuncSetAttribute");
        if (rc) return rc;
        attr_kern[dev][ki] = true;
    }
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    void* ws = nullptr;
    int rc = mlp_workspace(256, &ws, true);
    if (rc) return rc;
    uint32_t* counter = static_cast<uint32_t*>(ws);
    MlpPublish pub;
    if ((rc = mlp_publish(&pub))) return rc;
    // The counters start at zero: cleared by the last solve's read-back, or here after a new
    // allocation or a call that ended before its read-back. The calls of one device do not overlap
    // (device_call_lock), so the flag needs no lock of its own.
    bool& clear = g_mlp_counter[dev].clear;
    if (!clear) hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<float*>(counter), int64_t(2), 0.0f);
    clear = false;
    LearnedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.st.Xk = Xk;
    a.st.X0 = X;
    a.st.nsys = nsys;
    a.st.h = h;
    a.st.packed = reinterpret_cast<const __bf16*>(net->packed);
    a.st.params = net->params;
    a.st.nm = norm_of(*net);
    a.q = dev_params(*method, *options);
    gs_method plain;
    std::memset(&plain, 0, sizeof(plain));
    plain.linsolve = GS_LU;
    plain.factor_precision = GS_FACTOR_WORKING;
    plain.globalization = GS_STATIC;
    plain.refinement = GS_NO_REFINEMENT;
    plain.static_alpha = 1.0;
    a.qf = dev_params(plain, *options);
    a.S = S;
    a.counters = counter;
    const int64_t nchunk = (nsys + kChunk - 1) / kChunk;
    int64_t blocks = (nchunk + kWsCons - 1) / kWsCons;
    if (blocks > cus) blocks = cus;
    const dim3 grid{unsigned(blocks), 1, 1}, block{64 * kWsWaves, 1, 1};
    const bool refine = method->refinement == GS_REFINEMENT;
    bool timed = timing_begin(st);
    if (v) {
        if (refine) hipLaunchKernelGGL((learned_solve_kernel<true, true>), grid, block, kWsLds, st, a);
        else hipLaunchKernelGGL((learned_solve_kernel<true, false>), grid, block, kWsLds, st, a);
    } else {
        if (refine) hipLaunchKernelGGL((learned_solve_kernel<false, true>), grid, block, kWsLds, st, a);
        else hipLaunchKernelGGL((learned_solve_kernel<false, false>), grid, block, kWsLds, st, a);
    }
    timed = timing_mark(st, timed);
    if ((rc = check(hipGetLastError(), "fused learned solve launch failed"))) return rc;
    hipLaunchKernelGGL(publish_counters_kernel, dim3(1), dim3(64), 0, st, counter, pub.dev);
    if ((rc = check(hipGetLastError(), "counter read-back launch failed"))) return rc;
    if ((rc = check(hipStreamSynchronize(st), "hipStreamSynchronize"))) return rc;
    clear = true;
    timing_end(timed);
    const uint32_t host[2] = {reinterpret_cast<volatile uint32_t*>(pub.host)[0],
                              reinterpret_cast<volatile uint32_t*>(pub.host)[1]};
    if (host[1] != 0) return set_error(GS_EHIP, "the fused learned solve's hand-over between its waves timed out");
    if (fallbacks) *fallbacks = int64_t(host[0]);
    return GS_OK;
}

int gs_mlp_grad(const gs_mlp* net, const double* Xk, const double* X1, double h, const int64_t* idx, int32_t batch,
                float* grad, float* loss, void* stream) {
    if (!net || !net->params || !Xk || !X1 || !idx || !grad || !loss)
        return set_error(GS_EINVAL, "arguments must not be NULL");
    if (batch <= 0 || batch % 32 != 0 || batch > (1 << 20)) return set_error(GS_EINVAL, "batch must be a positive multiple of 32 up to 2^20");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t B = size_t(batch);
    // workspace: U (B×kIn bf16), T (B×kOut f32), Z1..Z3 and A1..A3 (B×kHid f32 / bf16), Z4 (B×kOut f32),
    // dZ4 (B×kOut f32), dA (B×kHid f32), dZ (B×kHid f32)
    // (12 regions, each rounded up to 256 bytes)
    const size_t bytes = B * kIn * 2 + B * kOut * 4 + 3 * B * kHid * 4 + 3 * B * kHid * 2 + 2 * B * kOut * 4 + 2 * B * kHid * 4 + 12 * 256;
    void* ws = nullptr;
    int rc = mlp_workspace(bytes, &ws);
    if (rc) return rc;
    char* p = static_cast<char*>(ws);
    auto take = [&](size_t n) {
        char* q = p;
        p += (n + 255) & ~size_t(255);
        return q;
    };
    __bf16* U = reinterpret_cast<__bf16*>(take(B * kIn * 2));
    float* T = reinterpret_cast<float*>(take(B * kOut * 4));
    float* Z[3];
    __bf16* A[3];
    for (int l = 0; l < 3; ++l) Z[l] = reinterpret_cast<float*>(take(B * kHid * 4));
    for (int l = 0; l < 3; ++l) A[l] = reinterpret_cast<__bf16*>(take(B * kHid * 2));
    float* Z4 = reinterpret_cast<float*>(take(B * kOut * 4));
    float* dZ4 = reinterpret_cast<float*>(take(B * kOut * 4));
    float* dA = reinterpret_cast<float*>(take(B * kHid * 4));
    float* dZ = reinterpret_cast<float*>(take(B * kHid * 4));
    const float* P = net->params;
    MlpNorm nm = norm_of(*net);
    hipLaunchKernelGGL(features_kernel, dim3(unsigned((B + 255) / 256)), dim3(256), 0, st, Xk, X1, idx, batch, h, nm, U, T);
    // forward
    {
        GemmArgs<__bf16, float> g{U, kIn, 1, P + kW1, kHid, 1, batch, kHid, kIn, kReluAct, P + kB1, Z[0], A[0], kHid};
        gemm(g, st);
    }
    for (int l = 1; l < 3; ++l) {
        const int wo = l == 1 ? kW2 : kW3, bo = l == 1 ? kB2 : kB3;
        GemmArgs<__bf16, float> g{A[l - 1], kHid, 1, P + wo, kHid, 1, batch, kHid, kHid, kReluAct, P + bo, Z[l], A[l], kHid};
        gemm(g, st);
    }
    {
        GemmArgs<__bf16, float> g{A[2], kHid, 1, P + kW4, kOut, 1, batch, kOut, kHid, kBias, P + kB4, Z4, nullptr, kOut};
        gemm(g, st);
    }
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, st, loss, int64_t(1), 0.0f);
    hipLaunchKernelGGL(loss_kernel, dim3(unsigned((B * kOut + 255) / 256)), dim3(256), 0, st, Z4, T, batch, dZ4, loss);
    // backward: the output layer
    {
        // dW4 = A3ᵀ dZ4 (kHid × kOut), db4 = Σ dZ4, dA3 = dZ4 W4ᵀ
        GemmArgs<__bf16, float> g{A[2], 1, kHid, dZ4, kOut, 1, kHid, kOut, batch, kPlain, nullptr, grad + kW4, nullptr, kOut};
        gemm(g, st);
        hipLaunchKernelGGL(colsum_kernel, dim3(kOut), dim3(256), 0, st, dZ4, batch, kOut, grad + kB4);
        GemmArgs<float, float> gd{dZ4, kOut, 1, P + kW4, 1, kOut, batch, kHid, kOut, kPlain, nullptr, dA, nullptr, kHid};
        gemm(gd, st);
    }
    for (int l = 2; l >= 0; --l) {
        hipLaunchKernelGGL(relu_backward_kernel, dim3(unsigned((B * kHid + 255) / 256)), dim3(256), 0, st, dA, Z[l],
                           int64_t(B * kHid), dZ);
        const int wo = l == 0 ? kW1 : l == 1 ? kW2 : kW3, bo = l == 0 ? kB1 : l == 1 ? kB2 : kB3;
        const int in = l == 0 ? kIn : kHid;
        if (l == 0) {
            GemmArgs<__bf16, float> g{U, 1, kIn, dZ, kHid, 1, kIn, kHid, batch, kPlain, nullptr, grad + wo, nullptr, kHid};
            gemm(g, st);
        } else {
            GemmArgs<__bf16, float> g{A[l - 1], 1, kHid, dZ, kHid, 1, in, kHid, batch, kPlain, nullptr, grad + wo, nullptr, kHid};
            gemm(g, st);
        }
        hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, st, dZ, batch, kHid, grad + bo);
        if (l > 0) {
            GemmArgs<float, float> gd{dZ, kHid, 1, P + wo, 1, kHid, batch, kHid, kHid, kPlain, nullptr, dA, nullptr, kHid};
            gemm(gd, st);
        }
    }
    return check(hipGetLastError(), "training launch failed");
}

int gs_mlp_adam(float* params, uint16_t* packed, const float* grad, float* m, float* v, int64_t step, float lr,
                float beta1, float beta2, float eps, void* stream) {
    if (!params || !packed || !grad || !m || !v) return set_error(GS_EINVAL, "arguments must not be NULL");
    if (step < 1) return set_error(GS_EINVAL, "step must be ≥ 1");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(adam_kernel, dim3(unsigned((kNParam + 255) / 256)), dim3(256), 0, st, params, grad, m, v, step, lr,
                       beta1, beta2, eps, reinterpret_cast<__bf16*>(packed));
    return check(hipGetLastError(), "Adam launch failed");
}

int gs_mlp_pack(const float* params, uint16_t* packed, void* stream) {
    if (!params || !packed) return set_error(GS_EINVAL, "arguments must not be NULL");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(pack_kernel, dim3(unsigned((kNParam + 255) / 256)), dim3(256), 0, st, params,
                       reinterpret_cast<__bf16*>(packed));
    return check(hipGetLastError(), "pack launch failed");
}

}  // extern "C"
