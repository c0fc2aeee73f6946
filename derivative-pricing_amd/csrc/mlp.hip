// Learned continuation value for Longstaff-Schwartz: a 2 -> 64 -> 64 -> 1 SiLU network,
// trained and evaluated with hand-written BF16 MFMA kernels (FP32 accumulation).
//
// No reference counterpart: the reference regresses on four Laguerre terms per exercise date
// (src/pricers/lsm.cpp:83-123). Here one network C(S/K, t/T) is fitted to the same
// regression targets (the discounted cash flows the QR backward induction produces at every
// ITM (path, date)), evaluated on a dense (S/K, date) grid, and the resulting table drives a
// streaming pricing kernel. QR coefficients remain the fallback outside the trained range
// and the judge of the learned exercise boundary.
//
// Layout: one wave owns 64 samples. The 64x64 layer and both backward products are
// 64x64x64 GEMMs C[m][n] = sum_k A[m][k] * Bt[n][k] on BF16 tiles in LDS (k contiguous,
// rows padded to 68 elements so a wave's 8-byte reads spread over all banks), computed as
// 2x2 tiles of v_mfma_f32_32x32x8_bf16. MFMA lane layout (32x32x8): lane l supplies
// A[l%32][4*(l/32)+j] and B[4*(l/32)+j][l%32], j < 4, and holds
// D[8*(r/4) + 4*(l/32) + r%4][l%32] in accumulator register r < 16.
#include <math.h>

#include <vector>

#include "mlp.hpp"
#include "runtime.hpp"

namespace mcdx {

constexpr int kH = 64;          // hidden width
constexpr int kLd = 68;         // padded LDS row (bf16 elements)
constexpr int kParams = 2 * kH + kH + kH * kH + kH + kH + 1; // 4417
static_assert(kParams == kMlpParams, "parameter count");
// parameter offsets in the flat FP32 vector
constexpr int oW1 = 0, oB1 = 2 * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH;

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Order this wave's LDS traffic (the wave executes in lockstep; this stops the compiler from
// moving one lane's reads of another lane's writes across the point).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned short f2bf(float f) { // round to nearest even
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}

__device__ __forceinline__ float silu(float a) { return a * fast_rcp(1.0f + fast_exp2(-a * 1.44269504f)); }
__device__ __forceinline__ float silu_grad(float a) {
    const float s = fast_rcp(1.0f + fast_exp2(-a * 1.44269504f));
    return s * (1.0f + a * (1.0f - s));
}

// acc[tm][tn] += A(rows tm*32.., k) * Bt(rows tn*32.., k) over k < 64.
__device__ __forceinline__ void wave_gemm64(const unsigned short* a, const unsigned short* bt,
                                            f16v acc[2][2]) {
    const int lane = threadIdx.x & 63;
    const int row = lane & 31, kg = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < kH; kk += 8) {
        s4v af[2], bf[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            af[t] = *(const s4v*)(a + (t * 32 + row) * kLd + kk + 4 * kg);
            bf[t] = *(const s4v*)(bt + (t * 32 + row) * kLd + kk + 4 * kg);
        }
#pragma unroll
        for (int tm = 0; tm < 2; ++tm)
#pragma unroll
            for (int tn = 0; tn < 2; ++tn)
                acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(af[tm], bf[tn], acc[tm][tn], 0, 0, 0);
    }
}

// Row (m) and column (n) of accumulator element r of tile (tm, tn) for this lane.
__device__ __forceinline__ int acc_row(int tm, int r) {
    const int lane = threadIdx.x & 63;
    return tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}
__device__ __forceinline__ int acc_col(int tn) { return tn * 32 + ((threadIdx.x & 63) & 31); }

struct MlpLds {
    unsigned short w2[kH * kLd];  // W2[o][i]
    unsigned short w2t[kH * kLd]; // W2[o][i] stored as [i][o]
    float w1[2 * kH], b1[kH], b2[kH], w3[kH], b3;
    float grad[2][kParams]; // per wave, summed in wave order by the update
    struct Wave {
        unsigned short h1[kH * kLd];  // H1[s][i]
        unsigned short h1t[kH * kLd]; // [i][s]
        unsigned short g2[kH * kLd];  // G2[s][o]; also FP32 H2 staging before backward
        unsigned short g2t[kH * kLd]; // [o][s]
        float xs[kH], ts[kH], g3[kH];
    } wave[2];
};

__device__ void load_weights(const float* __restrict__ w, MlpLds& L) {
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
        const int o = i / kH, k = i % kH;
        const unsigned short b = f2bf(w[oW2 + i]);
        L.w2[o * kLd + k] = b;
        L.w2t[k * kLd + o] = b;
    }
    for (int i = threadIdx.x; i < 2 * kH; i += blockDim.x) L.w1[i] = w[oW1 + i];
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        L.b1[i] = w[oB1 + i];
        L.b2[i] = w[oB2 + i];
        L.w3[i] = w[oW3 + i];
    }
    if (threadIdx.x == 0) L.b3 = w[oB3];
}

// Forward of this wave's 64 samples (x, t per lane); returns the network output of this
// lane's sample and leaves H1 / H1t in LDS and a2 (pre-activation of layer 2) in acc.
__device__ float forward64(MlpLds& L, MlpLds::Wave& W, float x, float t, f16v acc[2][2]) {
    const int s = threadIdx.x & 63;
    W.xs[s] = x;
    W.ts[s] = t;
    for (int i = 0; i < kH; ++i) {
        const float a1 = __builtin_fmaf(L.w1[2 * i], x, __builtin_fmaf(L.w1[2 * i + 1], t, L.b1[i]));
        const unsigned short h = f2bf(silu(a1));
        W.h1[s * kLd + i] = h;
        W.h1t[i * kLd + s] = h;
    }
    wave_sync();
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[tm][tn][r] = 0.f;
    wave_gemm64(W.h1, L.w2, acc);
    // a2 += b2; stage h2 * w3 per (s, o) in FP32 over the g2 area, then each lane sums its row
    float* h2w = (float*)W.g2; // [64][65] floats fits in g2 + g2t
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int o = acc_col(tn);
                acc[tm][tn][r] += L.b2[o];
                h2w[acc_row(tm, r) * 65 + o] = silu(acc[tm][tn][r]) * L.w3[o];
            }
    wave_sync();
    float y = L.b3;
    for (int o = 0; o < kH; ++o) y += h2w[s * 65 + o];
    wave_sync();
    return y;
}

// Backward of this wave's samples given dL/dy per lane; writes the wave's parameter
// gradients into L.grad[wave].
__device__ void backward64(MlpLds& L, MlpLds::Wave& W, float g3, f16v acc[2][2], int wave) {
    const int lane = threadIdx.x & 63;
    W.g3[lane] = g3;
    wave_sync();
    // G2 = g3 * w3 * silu'(a2); dW3 = sum_s g3 * h2; db2 = sum_s G2
    float dw3[2] = {0.f, 0.f}, db2[2] = {0.f, 0.f};
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int s = acc_row(tm, r), o = acc_col(tn);
                const float a2 = acc[tm][tn][r];
                const float gs = W.g3[s];
                dw3[tn] += gs * silu(a2);
                const float g2 = gs * L.w3[o] * silu_grad(a2);
                db2[tn] += g2;
                const unsigned short gb = f2bf(g2);
                W.g2[s * kLd + o] = gb;
                W.g2t[o * kLd + s] = gb;
            }
    // combine the two row halves (lanes l and l+32 share a column)
    float dw3o[2], db2o[2];
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        dw3o[tn] = dw3[tn] + __shfl_xor(dw3[tn], 32, 64);
        db2o[tn] = db2[tn] + __shfl_xor(db2[tn], 32, 64);
    }
    wave_sync();
    // dW2[o][i] = sum_s G2[s][o] H1[s][i]
    f16v dw2[2][2];
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) dw2[tm][tn][r] = 0.f;
    wave_gemm64(W.g2t, W.h1t, dw2);
    // G1[s][i] = sum_o G2[s][o] W2[o][i]
    f16v g1[2][2];
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) g1[tm][tn][r] = 0.f;
    wave_gemm64(W.g2, L.w2t, g1);
    float dw1a[2] = {0.f, 0.f}, dw1b[2] = {0.f, 0.f}, db1[2] = {0.f, 0.f};
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int s = acc_row(tm, r), i = acc_col(tn);
                const float x = W.xs[s], t = W.ts[s];
                const float a1 = __builtin_fmaf(L.w1[2 * i], x, __builtin_fmaf(L.w1[2 * i + 1], t, L.b1[i]));
                const float g = g1[tm][tn][r] * silu_grad(a1);
                dw1a[tn] += g * x;
                dw1b[tn] += g * t;
                db1[tn] += g;
            }
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        dw1a[tn] += __shfl_xor(dw1a[tn], 32, 64);
        dw1b[tn] += __shfl_xor(dw1b[tn], 32, 64);
        db1[tn] += __shfl_xor(db1[tn], 32, 64);
    }
    float gb3 = g3;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) gb3 += __shfl_xor(gb3, off, 64);
    float* gw = L.grad[wave];
    auto put = [&](int idx, float v) { gw[idx] = v; };
    if (lane < 32) {
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            const int c = tn * 32 + lane;
            put(oW3 + c, dw3o[tn]);
            put(oB2 + c, db2o[tn]);
            put(oW1 + 2 * c, dw1a[tn]);
            put(oW1 + 2 * c + 1, dw1b[tn]);
            put(oB1 + c, db1[tn]);
        }
    }
    if (lane == 0) put(oB3, gb3);
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int r = 0; r < 16; ++r) put(oW2 + acc_row(tm, r) * kH + acc_col(tn), dw2[tm][tn][r]);
}

struct TrainArgs {
    float* w;          // [kParams] weights (in/out)
    const float* xs;   // [n] first feature
    const float* ts;   // [n] second feature
    const float* ys;   // [n] targets (NaN = no sample)
    uint64_t n;
    int32_t steps;
    float lr, lr_decay;
    uint64_t seed;
    float* loss_out;   // [steps] batch losses
    float* grad_out;   // when non-null: one batch (indices 0..127), gradients written, no update
    int32_t avg_from;  // the weights returned are the mean of the iterates after steps
                       // avg_from .. steps - 1 (Polyak tail averaging; avg_from >= steps - 1:
                       // the last iterate)
};

constexpr int kTrainLanes = 128;
constexpr int kPer = (kParams + kTrainLanes - 1) / kTrainLanes; // parameters per lane (35)

// Lane-owned parameter i = lane + 128 k into its LDS places (BF16 W2 in both layouts).
__device__ __forceinline__ void scatter_param(MlpLds& L, int i, float v) {
    if (i < oB1) L.w1[i] = v;
    else if (i < oW2) L.b1[i - oB1] = v;
    else if (i < oB2) {
        const int j = i - oW2, o = j >> 6, k = j & 63;
        const unsigned short b = f2bf(v);
        L.w2[o * kLd + k] = b;
        L.w2t[k * kLd + o] = b;
    } else if (i < oW3) L.b2[i - oB2] = v;
    else if (i < oB3) L.w3[i - oW3] = v;
    else L.b3 = v;
}

// Single workgroup of 128 lanes (2 waves x 64 samples) running every Adam step in one
// launch. Each lane keeps its 35 parameters and their Adam moments in registers; the two
// waves' backward passes run concurrently into separate gradient arrays.
static_assert(sizeof(MlpLds) <= 160 * 1024, "LDS budget");

struct TrainLds {
    MlpLds net;
};

__global__ __launch_bounds__(kTrainLanes) void mlp_train_kernel(TrainArgs a) {
    extern __shared__ unsigned char smem[];
    TrainLds& T = *(TrainLds*)smem;
    MlpLds& L = T.net;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    MlpLds::Wave& W = L.wave[wave];
    const float b1c = 0.9f, b2c = 0.999f, eps = 1e-8f;
    float bc1 = 1.f, bc2 = 1.f, lr = a.lr;
    float wr[kPer], am[kPer], av[kPer]; // parameters and Adam moments, lane-owned
    float wa[kPer];                     // running mean of the iterates from avg_from on
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = tid + kTrainLanes * k;
        wr[k] = i < kParams ? a.w[i] : 0.f;
        am[k] = 0.f;
        av[k] = 0.f;
        wa[k] = 0.f;
    }
    for (int step = 0; step < a.steps; ++step) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + kTrainLanes * k;
            if (i < kParams) scatter_param(L, i, wr[k]);
        }
        __syncthreads();
        uint64_t idx = 0;
        float y = 0.f;
        if (a.grad_out) {
            idx = (uint64_t)tid < a.n ? (uint64_t)tid : 0;
            y = a.ys[idx];
        } else {
            // Rejection sampling over the valid (non-NaN target) samples; the draw index is
            // the attempt, so every lane's choice is a fixed function of (seed, step, lane).
            const uint64_t ctr = (uint64_t)step * kTrainLanes + tid;
            for (uint32_t tr = 0; tr < 8; ++tr) {
                uint64_t r = philox_bits(a.seed, ctr, tr, 7);
                if (a.n > 0xFFFFFFFFull) r |= (uint64_t)philox_bits(a.seed, ctr, tr + 8, 7) << 32;
                idx = r % a.n;
                y = a.ys[idx];
                if (y == y) break;
            }
        }
        const float valid = (y == y) ? 1.f : 0.f;
        const float x = valid != 0.f ? a.xs[idx] : 0.f, t = valid != 0.f ? a.ts[idx] : 0.f;
        if (valid == 0.f) y = 0.f;
        f16v acc[2][2];
        const float yhat = forward64(L, W, x, t, acc);
        const float diff = (yhat - y) * valid;
        const float g3 = 2.0f * diff / 128.0f;
        float l2 = diff * diff;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) l2 += __shfl_xor(l2, off, 64);
        backward64(L, W, g3, acc, wave);
        __syncthreads();
        if (a.grad_out) {
            for (int i = tid; i < kParams; i += kTrainLanes) a.grad_out[i] = L.grad[0][i] + L.grad[1][i];
            if (tid == 0 && a.loss_out) a.loss_out[0] = l2;
            if (tid == 64 && a.loss_out) a.loss_out[1] = l2;
            return;
        }
        if (tid == 0 && a.loss_out) a.loss_out[step] = l2 / 64.0f; // wave 0 batch half
        // Adam (FP32)
        bc1 *= b1c;
        bc2 *= b2c;
        // bias corrections once per step; the per-parameter square root and reciprocal at the
        // hardware rate (within an ulp or two of sqrtf and the IEEE division)
        const float c1 = 1.f / (1.f - bc1), c2 = 1.f / (1.f - bc2);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + kTrainLanes * k;
            if (i < kParams) {
                const float g = L.grad[0][i] + L.grad[1][i];
                am[k] = b1c * am[k] + (1.f - b1c) * g;
                av[k] = b2c * av[k] + (1.f - b2c) * g * g;
                const float mh = am[k] * c1, vh = av[k] * c2;
                wr[k] -= lr * mh * fast_rcp(fast_sqrt(vh) + eps);
            }
        }
        if (step >= a.avg_from) { // running mean of the iterates (uniform branch)
            const float inv = 1.f / (float)(step - a.avg_from + 1);
#pragma unroll
            for (int k = 0; k < kPer; ++k) wa[k] += (wr[k] - wa[k]) * inv;
        }
        lr *= a.lr_decay;
        __syncthreads(); // every lane's reads of the LDS weights and gradients are done
    }
    const bool averaged = a.avg_from < a.steps - 1;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = tid + kTrainLanes * k;
        if (i < kParams) a.w[i] = averaged ? wa[k] : wr[k];
    }
}

// Forward only, 64 samples per wave, any number of waves; out[i] = C(xs[i], ts[i]).
__global__ __launch_bounds__(128) void mlp_eval_kernel(const float* __restrict__ w,
                                                       const float* __restrict__ xs,
                                                       const float* __restrict__ ts, uint64_t n,
                                                       float* __restrict__ out) {
    extern __shared__ unsigned char smem[];
    MlpLds& L = *(MlpLds*)smem;
    load_weights(w, L);
    __syncthreads();
    MlpLds::Wave& W = L.wave[threadIdx.x >> 6];
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t j = i < n ? i : n - 1;
    f16v acc[2][2];
    const float y = forward64(L, W, xs[j], ts[j], acc);
    if (i < n) out[i] = y;
}

// w[i] from Philox(seed, i): layer-wise scales (see mlp.hpp).
__global__ __launch_bounds__(256) void mlp_init_kernel(float* w, uint64_t seed, float y_mean) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float z = inv_normal_f32(philox_w0((uint32_t)i, 0u, 0u, 9u, (uint32_t)seed,
                                             (uint32_t)(seed >> 32)));
    float val;
    if (i < oB1) val = z;                      // W1: unit-scale features
    else if (i < oW2) val = z;                 // b1
    else if (i < oB2) val = z * 0.125f * 1.41421356f; // W2: He scale sqrt(2/64)
    else if (i < oW3) val = 0.f;               // b2
    else if (i < oB3) val = z * 0.125f;        // W3
    else val = y_mean;                         // b3
    w[i] = val;
}

namespace {
size_t lds_bytes() { return sizeof(MlpLds); }
} // namespace

int mlp_init_device(float* w, uint64_t seed, float y_mean, hipStream_t s) {
    hipLaunchKernelGGL(mlp_init_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, w, seed,
                       y_mean);
    return status_of(launch_check());
}

int mlp_train_device(float* w, const float* xs, const float* ts, const float* ys, uint64_t n,
                     int steps, float lr, float lr_decay, uint64_t seed, float* loss, float* grad,
                     hipStream_t s, int avg_from) {
    TrainArgs a{w, xs, ts, ys, n, steps, lr, lr_decay, seed, loss, grad,
                avg_from < 0 ? steps : avg_from};
    const int bytes = (int)sizeof(TrainLds);
    hipError_t e = hipFuncSetAttribute((const void*)mlp_train_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kTrainLanes), bytes, s, a);
    return status_of(launch_check());
}

int mlp_eval_device(const float* w, const float* xs, const float* ts, uint64_t n, float* out,
                    hipStream_t s) {
    if (n == 0) return MCDX_OK;
    hipError_t e = hipFuncSetAttribute((const void*)mlp_eval_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes());
    if (e != hipSuccess) return status_of(e);
    hipLaunchKernelGGL(mlp_eval_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), lds_bytes(),
                       s, w, xs, ts, n, out);
    return status_of(launch_check());
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

uint32_t mcdx_mlp_param_count(void) { return kParams; }

// Gradient of the batch-mean squared error (weights FP32; BF16 MFMA products, FP32
// accumulation) for n <= 128 samples (the batch is padded by repeating sample 0 when n <
// 128; callers pass exactly 128 to compare with an exact reference). loss2[0..1] receive the
// two waves' summed squared errors.
int mcdx_mlp_gradients(const float* weights, const float* xs, const float* ts, const float* ys,
                       uint32_t n, float* grad, float* loss2) {
    if (!weights || !xs || !ts || !ys || !grad || n == 0 || n > 128) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const size_t pb = kParams * sizeof(float);
    const size_t bytes = 3 * pb + 3 * 128 * sizeof(float) + pb + 64;
    if ((st = ensure_dev(ctx, bytes))) return st;
    char* base = (char*)ctx->dev;
    float* dw = (float*)base;
    float* dm = dw + kParams;
    float* dv = dm + kParams;
    float* dx = dv + kParams;
    float* dt = dx + 128;
    float* dy = dt + 128;
    float* dg = dy + 128;
    float* dl = dg + kParams;
    hipError_t e = hipMemcpy(dw, weights, pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dx, xs, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, ts, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, ys, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return status_of(e);
    if ((st = mlp_train_device(dw, dx, dt, dy, n, 1, 0.f, 1.f, 0, dl, dg, ctx->stream))) return st;
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(grad, dg, pb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss2) e = hipMemcpy(loss2, dl, 8, hipMemcpyDeviceToHost);
    return status_of(e);
}

// `steps` Adam steps from `weights` (updated in place) with zero initial moments; samples with
// a NaN target are skipped. losses[step] (optional) = first-half batch MSE per step.
int mcdx_mlp_train(float* weights, const float* xs, const float* ts, const float* ys,
                   uint64_t n, int32_t steps, float learning_rate, float lr_decay,
                   uint64_t seed, float* losses) {
    if (!weights || !xs || !ts || !ys || n == 0 || steps < 1 || n > (1ull << 28) ||
        steps > (1 << 24))
        return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const size_t pb = kParams * sizeof(float);
    const size_t bytes = 3 * pb + 3 * n * sizeof(float) + (size_t)steps * sizeof(float) + 1024;
    if ((st = ensure_dev(ctx, bytes))) return st;
    float* dw = (float*)ctx->dev;
    float* dm = dw + kParams;
    float* dv = dm + kParams;
    float* dx = dv + kParams;
    float* dt = dx + n;
    float* dy = dt + n;
    float* dl = dy + n;
    hipError_t e = hipMemcpy(dw, weights, pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dx, xs, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, ts, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, ys, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return status_of(e);
    if ((st = mlp_train_device(dw, dx, dt, dy, n, steps, learning_rate, lr_decay, seed, dl,
                               nullptr, ctx->stream)))
        return st;
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(weights, dw, pb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && losses) e = hipMemcpy(losses, dl, (size_t)steps * 4, hipMemcpyDeviceToHost);
    return status_of(e);
}

// out[i] = network(xs[i], ts[i]) through the MFMA forward kernel.
int mcdx_mlp_forward(const float* weights, const float* xs, const float* ts, uint64_t n,
                     float* out) {
    if (!weights || (!xs && n) || (!ts && n) || (!out && n)) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 26)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const size_t pb = kParams * sizeof(float);
    if ((st = ensure_dev(ctx, pb + 3 * n * sizeof(float) + 256))) return st;
    float* dw = (float*)ctx->dev;
    float* dx = dw + kParams;
    float* dt = dx + n;
    float* dout = dt + n;
    hipError_t e = hipMemcpy(dw, weights, pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dx, xs, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, ts, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return status_of(e);
    if ((st = mlp_eval_device(dw, dx, dt, n, dout, ctx->stream))) return st;
    e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost);
    return status_of(e);
}

} // extern "C"
