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
// Layout: the 64x64 layer and both backward products are 64x64x64 GEMMs
// C[m][n] = sum_k A[m][k] * Bt[n][k] on BF16 tiles in LDS (k contiguous, rows padded to 72
// elements: 16-byte aligned, and the 16-byte reads of 16 consecutive lanes land on distinct
// banks), computed as 32x32 tiles of gfx950's v_mfma_f32_32x32x16_bf16. MFMA lane layout
// (32x32x16): lane l supplies A[l%32][8*(l/32)+j] and B[8*(l/32)+j][l%32], j < 8, and holds
// D[8*(r/4) + 4*(l/32) + r%4][l%32] in accumulator register r < 16.
#include <math.h>

#include <vector>

#include "mlp.hpp"
#include "runtime.hpp"

namespace mcdx {

constexpr int kH = 64;          // hidden width
constexpr int kLd = 72;         // padded LDS row (bf16 elements)
constexpr int kParams = 2 * kH + kH + kH * kH + kH + kH + 1; // 4417
static_assert(kParams == kMlpParams, "parameter count");
// parameter offsets in the flat FP32 vector
constexpr int oW1 = 0, oB1 = 2 * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH;

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Order this wave's LDS traffic (the wave executes in lockstep; this stops the compiler from
// moving one lane's reads of another lane's writes across the point).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned short f2bf(float f) {
    return __builtin_bit_cast(unsigned short, (__bf16)f);
}
// Two values in one conversion: lo in bits 0-15, hi in bits 16-31.
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    const bf2v v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ float sigmoid(float a) { return fast_rcp(1.0f + fast_exp2(-a * 1.44269504f)); }
__device__ __forceinline__ float silu(float a) { return a * sigmoid(a); }
__device__ __forceinline__ float silu_grad(float a) {
    const float s = sigmoid(a);
    return s * (1.0f + a * (1.0f - s));
}

// Row (m) and column (n) of accumulator element r of tile (tm, tn) for this lane.
__device__ __forceinline__ int acc_row(int tm, int r) {
    const int lane = threadIdx.x & 63;
    return tm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}
__device__ __forceinline__ int acc_col(int tn) { return tn * 32 + ((threadIdx.x & 63) & 31); }

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

// Training layout: one workgroup of 8 waves. A step's batch is two blocks of 64 samples (lane
// = sample); each block's four waves split its work by accumulator tile: wave q of a block
// owns tile (tm, tn) = (q / 2, q % 2) of every 64x64 product (one 32x32x16 MFMA chain instead
// of four) and the elementwise work on that tile, and hidden units 16q .. 16q+15 of the first
// layer. Column sums over a tile's rows are combined across the two row tiles through LDS.
// A step is seven barrier-separated phases: first layer; second layer and output staging;
// output and dL/dy; output-layer backward; dW2 and the first-layer backward; column sums;
// Adam with the next step's weights and batch written to LDS.
constexpr int kSampleBlocks = 2;
constexpr int kWavesPerBlock = 4;
constexpr int kTrainLanes = 64 * kSampleBlocks * kWavesPerBlock; // 512
constexpr int kPer = (kParams + kTrainLanes - 1) / kTrainLanes;  // parameters per lane (9)
constexpr int kAttempts = 8; // rejection-sampling attempts per sample (two Philox calls)

// Single-wave roles, placed so that the two blocks' busy waves of a phase sit on different
// SIMDs (wave w of the workgroup runs on SIMD w % 4): the output wave sums each sample's
// output-layer row, the sampler wave draws the next step's batch, the column wave combines
// the row tiles' column sums.
__device__ __forceinline__ int output_wave(int bi) { return bi ? 2 : 0; }
__device__ __forceinline__ int sampler_wave(int bi) { return bi ? 1 : 3; }
__device__ __forceinline__ int column_wave(int bi) { return bi ? 3 : 1; }

// All four words of Philox4x32-10({c0,c1,c2,c3}, {k0,k1}) (the same rounds as philox_w0).
__device__ __forceinline__ void philox4(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c[0];
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c[2];
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c[1], k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c[3], k1);
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
}

// a mod n for 32-bit a and n >= 1 without a division: M = floor((2^64 - 1) / n) + 1 (Lemire,
// Kaser and Kurz, "Faster remainder by direct computation", 2019).
__device__ __forceinline__ uint32_t fastmod_u32(uint32_t a, uint64_t m, uint32_t n) {
    return (uint32_t)__umul64hi(m * a, (uint64_t)n);
}

struct TrainLds {
    alignas(16) unsigned short w2[kH * kLd];  // W2[o][i]
    alignas(16) unsigned short w2t[kH * kLd]; // W2[o][i] stored as [i][o]
    alignas(16) float w1[2 * kH];
    alignas(16) float b1[kH];
    float b2[kH], w3[kH], b3;
    float grad[kSampleBlocks][kParams]; // per sample block, summed in block order by Adam
    struct Block {
        alignas(16) unsigned short h1[kH * kLd];  // H1[s][i]
        alignas(16) unsigned short h1t[kH * kLd]; // [i][s]
        alignas(16) unsigned short g2[kH * kLd];  // G2[s][o]; FP32 h2 * w3 staging before the backward
        alignas(16) unsigned short g2t[kH * kLd]; // [o][s]
        alignas(16) float xs[kH];
        alignas(16) float ts[kH];
        alignas(16) float ys[kH];
        alignas(16) float g3[kH];
        float part[5][2][kH]; // dW3, db2, dW1[:,0], dW1[:,1], db1 column sums per row tile
    } blk[kSampleBlocks];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "LDS budget");

// acc += A(rows tm*32.., k) * Bt(rows tn*32.., k) over k < 64: one 32x32 tile.
__device__ __forceinline__ void tile_gemm64(const unsigned short* a, const unsigned short* bt,
                                            int tm, int tn, f16v& acc) {
    const int lane = threadIdx.x & 63;
    const int row = lane & 31, kg = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < kH; kk += 16) {
        const bf8v af = *(const bf8v*)(a + (tm * 32 + row) * kLd + kk + 8 * kg);
        const bf8v bf = *(const bf8v*)(bt + (tn * 32 + row) * kLd + kk + 8 * kg);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, bf, acc, 0, 0, 0);
    }
}

// Lane-owned parameter i = lane + kTrainLanes k into its LDS places (BF16 W2 in both layouts).
__device__ __forceinline__ void scatter_param(TrainLds& L, int i, float v) {
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

// Every Adam step in one launch. Each lane keeps its kPer parameters and their Adam moments
// in registers; per step the batch's gradient is formed in LDS by the eight waves.
__global__ __launch_bounds__(kTrainLanes) void mlp_train_kernel(TrainArgs a) {
    extern __shared__ unsigned char smem[];
    TrainLds& L = *(TrainLds*)smem;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int bi = tid >> 8;                       // sample block
    const int q = (tid >> 6) & (kWavesPerBlock - 1); // wave within the block
    const int tm = q >> 1, tn = q & 1;
    TrainLds::Block& B = L.blk[bi];
    float* gw = L.grad[bi];
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
    // Sampling: rejection sampling over the valid (non-NaN target) samples, up to kAttempts
    // attempts, so every sample's choice is a fixed function of (seed, step, sample); after
    // kAttempts invalid draws the sample is masked. Attempt j of sample c at step st is word
    // j % 4 of Philox(key seed, counter {c, 0, j / 4, 7}) with c = st * 128 + sample, reduced
    // mod n (beyond 2^32 samples: words j of two calls make 64 bits). The sampler wave of each
    // block prefetches: it issues all attempts' target loads for the next step while the
    // output wave works, picks the first valid attempt and issues its feature loads while the
    // column wave works, and stores the sample at the next step's start, so no HBM round trip
    // sits on the step's critical path.
    const int sample = bi * 64 + lane;
    const int q_out = output_wave(bi), q_smp = sampler_wave(bi), q_col = column_wave(bi);
    uint64_t pidx[kAttempts];
    float pty[kAttempts];
    float nx = 0.f, nt = 0.f, ny = 0.f;
    const bool wide = a.n > 0xFFFFFFFFull; // uniform
    const uint64_t fm = wide ? 0 : 0xFFFFFFFFFFFFFFFFull / a.n + 1;
    auto issue = [&](int st) {
        const uint64_t ctr = (uint64_t)st * 128 + sample;
#pragma unroll
        for (uint32_t h = 0; h < kAttempts / 4; ++h) {
            uint32_t w[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), h, 7u};
            philox4(w, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
            uint32_t v[4] = {0u, 0u, 0u, 0u};
            if (wide) {
                v[0] = (uint32_t)ctr;
                v[1] = (uint32_t)(ctr >> 32);
                v[2] = h + kAttempts / 4;
                v[3] = 7u;
                philox4(v, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int tr = 4 * h + j;
                pidx[tr] = wide ? (((uint64_t)v[j] << 32) | w[j]) % a.n
                                : fastmod_u32(w[j], fm, (uint32_t)a.n);
                pty[tr] = a.ys[pidx[tr]];
            }
        }
    };
    auto choose = [&]() {
        uint64_t idx = pidx[kAttempts - 1];
        float y = pty[kAttempts - 1];
#pragma unroll
        for (int tr = kAttempts - 1; tr >= 0; --tr)
            if (pty[tr] == pty[tr]) {
                idx = pidx[tr];
                y = pty[tr];
            }
        const bool valid = y == y;
        nx = valid ? a.xs[idx] : 0.f;
        nt = valid ? a.ts[idx] : 0.f;
        ny = y;
    };
    // The weights and the batch of step 0; later steps get theirs from the previous step's
    // update phase.
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = tid + kTrainLanes * k;
        if (i < kParams) scatter_param(L, i, wr[k]);
    }
    // sample s of block bi is the batch's sample bi * 64 + s
    if (a.grad_out) {
        if (q == 0) {
            const uint64_t idx = (uint64_t)sample < a.n ? (uint64_t)sample : 0;
            const float y = a.ys[idx];
            const bool valid = y == y;
            B.xs[lane] = valid ? a.xs[idx] : 0.f;
            B.ts[lane] = valid ? a.ts[idx] : 0.f;
            B.ys[lane] = y;
        }
    } else if (q == q_smp) {
        issue(0);
        choose();
        B.xs[lane] = nx;
        B.ts[lane] = nt;
        B.ys[lane] = ny;
    }
    __syncthreads();
    for (int step = 0; step < a.steps; ++step) {
        // first layer: hidden units 16q .. 16q+15 of this lane's sample
        {
            const float x = B.xs[lane], t = B.ts[lane];
            // the wave's 16 units: weights by 16-byte broadcast reads, the lane's 16 values
            // of H1 as two 16-byte stores (its row segment is contiguous)
            const float4* w1v = (const float4*)(L.w1 + 32 * q);
            const float4* b1v = (const float4*)(L.b1 + 16 * q);
            uint32_t packed[8];
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) {
                const float4 wa = w1v[2 * u4], wb = w1v[2 * u4 + 1], bb = b1v[u4];
                const float w[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
                const float b[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
                for (int j = 0; j < 4; j += 2) {
                    const int u = 4 * u4 + j, i = q * 16 + u;
                    const float a0 = __builtin_fmaf(w[2 * j], x, __builtin_fmaf(w[2 * j + 1], t, b[j]));
                    const float a1 = __builtin_fmaf(w[2 * j + 2], x, __builtin_fmaf(w[2 * j + 3], t, b[j + 1]));
                    const uint32_t h = f2bf2(silu(a0), silu(a1));
                    B.h1t[i * kLd + lane] = (unsigned short)h;
                    B.h1t[(i + 1) * kLd + lane] = (unsigned short)(h >> 16);
                    packed[u >> 1] = h;
                }
            }
            uint4* row = (uint4*)(B.h1 + lane * kLd + 16 * q);
            row[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
            row[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
        }
        __syncthreads();
        // second layer on this wave's tile; stage h2 * w3 per (s, o) in FP32
        f16v acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
        tile_gemm64(B.h1, L.w2, tm, tn, acc);
        float* h2w = (float*)B.g2; // [64][65] floats fits in g2 + g2t
        {
            const int o = acc_col(tn);
            const float b2o = L.b2[o], w3o = L.w3[o];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc[r] += b2o;
                h2w[acc_row(tm, r) * 65 + o] = silu(acc[r]) * w3o;
            }
        }
        __syncthreads();
        float l2 = 0.f;
        if (!a.grad_out && q == q_smp && step + 1 < a.steps) issue(step + 1);
        if (q == q_out) { // output and dL/dy of this lane's sample (row sum in column order)
            float yhat = L.b3;
            for (int o = 0; o < kH; ++o) yhat += h2w[lane * 65 + o];
            const float y = B.ys[lane];
            const float diff = (y == y) ? (yhat - y) : 0.f;
            B.g3[lane] = 2.0f * diff / 128.0f;
            l2 = diff * diff;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) l2 += __shfl_xor(l2, off, 64);
        }
        __syncthreads();
        // backward through the output layer on this wave's tile: G2 = g3 * w3 * silu'(a2)
        {
            const int o = acc_col(tn);
            const float w3o = L.w3[o];
            float dw3 = 0.f, db2 = 0.f;
            // rows come in runs of four consecutive samples (accumulator registers 4rg..4rg+3):
            // their dL/dy by one 16-byte read, their G2^T entries by one 8-byte store
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const int s0 = acc_row(tm, 4 * rg);
                const float4 g3v = *(const float4*)(B.g3 + s0);
                const float gsv[4] = {g3v.x, g3v.y, g3v.z, g3v.w};
                float g2v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float a2 = acc[4 * rg + j];
                    const float gs = gsv[j];
                    const float sg = sigmoid(a2); // silu and its derivative share it
                    dw3 += gs * (a2 * sg);
                    g2v[j] = gs * w3o * (sg * (1.0f + a2 * (1.0f - sg)));
                    db2 += g2v[j];
                }
                const uint32_t p01 = f2bf2(g2v[0], g2v[1]), p23 = f2bf2(g2v[2], g2v[3]);
                B.g2[s0 * kLd + o] = (unsigned short)p01;
                B.g2[(s0 + 1) * kLd + o] = (unsigned short)(p01 >> 16);
                B.g2[(s0 + 2) * kLd + o] = (unsigned short)p23;
                B.g2[(s0 + 3) * kLd + o] = (unsigned short)(p23 >> 16);
                *(uint2*)(B.g2t + o * kLd + s0) = make_uint2(p01, p23);
            }
            dw3 += __shfl_xor(dw3, 32, 64);
            db2 += __shfl_xor(db2, 32, 64);
            if (lane < 32) {
                B.part[0][tm][o] = dw3;
                B.part[1][tm][o] = db2;
            }
        }
        __syncthreads();
        // dW2[o][i] = sum_s G2[s][o] H1[s][i] on this wave's tile
        {
            f16v dw2;
#pragma unroll
            for (int r = 0; r < 16; ++r) dw2[r] = 0.f;
            tile_gemm64(B.g2t, B.h1t, tm, tn, dw2);
#pragma unroll
            for (int r = 0; r < 16; ++r) gw[oW2 + acc_row(tm, r) * kH + acc_col(tn)] = dw2[r];
        }
        // G1[s][i] = sum_o G2[s][o] W2[o][i], then through the first layer's SiLU
        {
            f16v g1;
#pragma unroll
            for (int r = 0; r < 16; ++r) g1[r] = 0.f;
            tile_gemm64(B.g2, L.w2t, tm, tn, g1);
            const int i = acc_col(tn);
            const float w1a = L.w1[2 * i], w1b = L.w1[2 * i + 1], b1i = L.b1[i];
            float dw1a = 0.f, dw1b = 0.f, db1 = 0.f;
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const int s0 = acc_row(tm, 4 * rg);
                const float4 xv = *(const float4*)(B.xs + s0), tv = *(const float4*)(B.ts + s0);
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w}, ta[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float x = xa[j], t = ta[j];
                    const float a1 = __builtin_fmaf(w1a, x, __builtin_fmaf(w1b, t, b1i));
                    const float g = g1[4 * rg + j] * silu_grad(a1);
                    dw1a += g * x;
                    dw1b += g * t;
                    db1 += g;
                }
            }
            dw1a += __shfl_xor(dw1a, 32, 64);
            dw1b += __shfl_xor(dw1b, 32, 64);
            db1 += __shfl_xor(db1, 32, 64);
            if (lane < 32) {
                B.part[2][tm][i] = dw1a;
                B.part[3][tm][i] = dw1b;
                B.part[4][tm][i] = db1;
            }
        }
        if (q == q_out) {
            float gb3 = B.g3[lane];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) gb3 += __shfl_xor(gb3, off, 64);
            if (lane == 0) gw[oB3] = gb3;
        }
        __syncthreads();
        if (!a.grad_out && q == q_smp && step + 1 < a.steps) choose();
        if (q == q_col) { // column sums of the two row tiles
            const int c = lane;
            gw[oW3 + c] = B.part[0][0][c] + B.part[0][1][c];
            gw[oB2 + c] = B.part[1][0][c] + B.part[1][1][c];
            gw[oW1 + 2 * c] = B.part[2][0][c] + B.part[2][1][c];
            gw[oW1 + 2 * c + 1] = B.part[3][0][c] + B.part[3][1][c];
            gw[oB1 + c] = B.part[4][0][c] + B.part[4][1][c];
        }
        __syncthreads();
        if (a.grad_out) {
            for (int i = tid; i < kParams; i += kTrainLanes) a.grad_out[i] = L.grad[0][i] + L.grad[1][i];
            if (q == q_out && lane == 0 && a.loss_out) a.loss_out[bi] = l2; // each block's sum
            return;
        }
        if (bi == 0 && q == q_out && lane == 0 && a.loss_out) a.loss_out[step] = l2 / 64.0f; // first block
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
        // the next step's weights and batch: every read of the LDS weights and of the batch
        // ended before the previous barrier
        if (step + 1 < a.steps) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int i = tid + kTrainLanes * k;
                if (i < kParams) scatter_param(L, i, wr[k]);
            }
            if (q == q_smp) { // chosen and loaded during this step
                B.xs[lane] = nx;
                B.ts[lane] = nt;
                B.ys[lane] = ny;
            }
        }
        __syncthreads(); // the next step's weights and batch are in place; gradients read
    }
    const bool averaged = a.avg_from < a.steps - 1;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = tid + kTrainLanes * k;
        if (i < kParams) a.w[i] = averaged ? wa[k] : wr[k];
    }
}

// Forward only: out[i] = C(x_i, t_i) for i < n. A block of four waves evaluates one 64-sample
// chunk (lane = sample), split as the trainer splits a step: wave q forms hidden units
// 16q .. 16q+15 of the first layer (weights read through the scalar unit), then accumulator
// tile (tm, tn) = (q / 2, q % 2) of the 64x64x64 layer-2 product on MFMA (A = the BF16 H1
// rows in LDS, B = W2 rows read from memory and rounded to BF16 as the trainer rounds them,
// FP32 accumulation over k in the trainer's order), SiLU and the output-layer products; wave 0
// then sums each sample's 64 products in a fixed order. The features come from xs, ts, or,
// with g.grid > 0, from the exercise-date grid: i = (d - 1) grid + k gives
// (fs (xlo + h k - 1), 2 d / m - 1).
__global__ __launch_bounds__(256) void mlp_eval_kernel(const float* __restrict__ w,
                                                       const float* __restrict__ xs,
                                                       const float* __restrict__ ts, uint64_t n,
                                                       MlpEvalGrid g, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) unsigned short h1[kH * kLd]; // H1[s][i]
    __shared__ float h2w[kH * (kH + 1)]; // h2[s][o] * w3[o], rows padded to 65
    const uint32_t q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;
    const uint64_t j = i < n ? i : n - 1;
    float x, t;
    if (g.grid) {
        const int date = (int)(j / g.grid) + 1;
        const uint32_t k = (uint32_t)(j % g.grid);
        x = (float)(g.feat_scale * (g.xlo + g.h * (double)k - 1.0));
        t = (float)(2.0 * (double)date / (double)g.m - 1.0);
    } else {
        x = xs[j];
        t = ts[j];
    }
#pragma unroll
    for (int u = 0; u < 16; u += 2) {
        const int i0 = 16 * (int)q + u;
        const float a0 = __builtin_fmaf(w[oW1 + 2 * i0], x, __builtin_fmaf(w[oW1 + 2 * i0 + 1], t, w[oB1 + i0]));
        const float a1 = __builtin_fmaf(w[oW1 + 2 * i0 + 2], x,
                                        __builtin_fmaf(w[oW1 + 2 * i0 + 3], t, w[oB1 + i0 + 1]));
        *(uint32_t*)(h1 + lane * kLd + i0) = (uint32_t)f2bf(silu(a0)) | ((uint32_t)f2bf(silu(a1)) << 16);
    }
    __syncthreads();
    const int tm = (int)q >> 1, tn = (int)q & 1;
    const int row = lane & 31, kg = lane >> 5;
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int kk = 0; kk < kH; kk += 16) {
        const bf8v af = *(const bf8v*)(h1 + (tm * 32 + row) * kLd + kk + 8 * kg);
        const float* wr = w + oW2 + (tn * 32 + row) * kH + kk + 8 * kg; // W2[o][k .. k+7]
        bf8v bfr;
#pragma unroll
        for (int e = 0; e < 8; ++e) bfr[e] = __builtin_bit_cast(__bf16, f2bf(wr[e]));
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, bfr, acc, 0, 0, 0);
    }
    const int o = acc_col(tn);
    const float b2 = w[oB2 + o], w3 = w[oW3 + o];
#pragma unroll
    for (int r = 0; r < 16; ++r) h2w[acc_row(tm, r) * (kH + 1) + o] = silu(acc[r] + b2) * w3;
    __syncthreads();
    if (q == 0) {
        float y = w[oB3];
        for (int c = 0; c < kH; ++c) y += h2w[lane * (kH + 1) + c];
        if (i < n) out[i] = y;
    }
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
                    hipStream_t s, const MlpEvalGrid* grid) {
    if (n == 0) return MCDX_OK;
    const uint64_t chunks = (n + 63) / 64; // one block per 64 samples
    if (chunks > 0x7FFFFFFFull) return MCDX_EINVAL;
    const MlpEvalGrid g = grid ? *grid : MlpEvalGrid{0, 0, 0.0, 0.0, 0.0};
    hipLaunchKernelGGL(mlp_eval_kernel, dim3((unsigned)chunks), dim3(256), 0, s, w, xs, ts, n, g, out);
    return status_of(launch_check());
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

uint32_t mcdx_mlp_param_count(void) { return kParams; }

// Gradient of the batch-mean squared error (weights FP32; BF16 MFMA products, FP32
// accumulation) for n <= 128 samples (the batch is padded by repeating sample 0 when n <
// 128; callers pass exactly 128 to compare with an exact reference). loss2[0..1] receive the
// two 64-sample blocks' summed squared errors.
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
