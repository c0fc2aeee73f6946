// Device primitives shared by every path kernel (gfx950, wave64).
//
//  * Philox4x32-10 in registers, one counter per lane: key {seed_lo, seed_hi},
//    counter {path_lo, path_hi, draw, stream}, output word 0 only
//    (reference src/core/rng.cpp:20-74). The key schedule and every counter word
//    that is uniform across the wave (draw, stream) fold onto the scalar unit.
//  * Inverse normal CDF, two precisions:
//      - FP64: Acklam + one Halley step against 0.5*erfc(-x/sqrt2), the reference's
//        exact arithmetic (src/core/normal.cpp:50-82), built without FP contraction.
//      - FP32: Acklam alone. Its 1.15e-9 relative error is below FP32 rounding, so the
//        Halley step cannot improve an FP32 result. The central/tail branch is decided
//        on the integer bits with thresholds that reproduce the reference's FP64
//        comparison exactly, and the tail argument is formed from min(bits, ~bits) so
//        1-u is exact (no u rounding to 1.0f).
//  * FP64 Welford accumulators and the Chan-Golub-LeVeque merge
//    (include/mcd/core/stats.hpp:21-65), merged along a fixed pairwise tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcdx {

// ---------------------------------------------------------------------------
// Philox4x32-10 (output word 0)
// ---------------------------------------------------------------------------
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// a ^ b ^ c in one VALU instruction on gfx950 (v_bitop3_b32, truth table 0x96).
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Word 0 of Philox4x32-10({c0,c1,c2,c3}, {k0,k1}); rounds as rng.cpp:26-30, key bump :32.
// Each 32x32->64 product is one v_mad_u64_u32 (both halves at once) and each three-way
// XOR one v_bitop3_b32: about 36 VALU instructions per draw once the compiler drops the
// words that word 0 does not depend on and folds the wave-uniform draw/stream words and
// the key schedule onto the scalar unit.
__host__ __device__ __forceinline__ uint32_t philox_w0(uint32_t c0, uint32_t c1, uint32_t c2,
                                                       uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return c0;
}

__host__ __device__ __forceinline__ uint32_t philox_bits(uint64_t seed, uint64_t path,
                                                         uint32_t draw, uint32_t stream) {
    return philox_w0((uint32_t)path, (uint32_t)(path >> 32), draw, stream, (uint32_t)seed,
                     (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------------------
// Inverse normal CDF
// ---------------------------------------------------------------------------
// Acklam coefficients (published constants; reference src/core/normal.cpp:22-45).
#define MCDX_ACKLAM_CENTRAL(T, S)                                                           \
    constexpr T A1 = S(-3.969683028665376e+01), A2 = S(2.209460984245205e+02),              \
                A3 = S(-2.759285104469687e+02), A4 = S(1.383577518672690e+02),              \
                A5 = S(-3.066479806614716e+01), A6 = S(2.506628277459239e+00);              \
    constexpr T B1 = S(-5.447609879822406e+01), B2 = S(1.615858368580409e+02),              \
                B3 = S(-1.556989798598866e+02), B4 = S(6.680131188771972e+01),              \
                B5 = S(-1.328068155288572e+01);
#define MCDX_ACKLAM_TAIL(T, S)                                                              \
    constexpr T C1 = S(-7.784894002430293e-03), C2 = S(-3.223964580411365e-01),             \
                C3 = S(-2.400758277161838e+00), C4 = S(-2.549732539343734e+00),             \
                C5 = S(4.374664141464968e+00), C6 = S(2.938163982698783e+00);               \
    constexpr T D1 = S(7.784695709041462e-03), D2 = S(3.224671290700398e-01),               \
                D3 = S(2.445134137142996e+00), D4 = S(3.754408661907416e+00);

#define MCDX_F32(x) ((float)(x))
#define MCDX_F64(x) ((double)(x))

// Branch thresholds on the raw bits: bits < kTailLo  <=>  (bits+0.5)/2^32 < 0.02425, and
// bits > kTailHi  <=>  (bits+0.5)/2^32 > 0.97575, both exactly as evaluated in FP64
// (0.02425*2^32 = 104152956.928, (1-0.02425)*2^32 = 4190814339.072 in FP64).
constexpr uint32_t kTailLo = 104152957u;  // first bits value with u >= 0.02425
constexpr uint32_t kTailHi = 4190814338u; // last bits value with u <= 1 - 0.02425

// Hardware-rate FP32 transcendentals (v_rcp_f32, v_log_f32, v_sqrt_f32, v_exp_f32, each
// within about 1 ulp over the ranges used here: arguments are normal floats, never denormal).
#if defined(__HIP_DEVICE_COMPILE__)
#define MCDX_DEV_OR_HOST(dev, host) dev
#else
#define MCDX_DEV_OR_HOST(dev, host) host
#endif
__host__ __device__ __forceinline__ float fast_rcp(float x) {
    return MCDX_DEV_OR_HOST(__builtin_amdgcn_rcpf(x), 1.0f / x);
}
__host__ __device__ __forceinline__ float fast_log2(float x) {
    return MCDX_DEV_OR_HOST(__builtin_amdgcn_logf(x), log2f(x));
}
__host__ __device__ __forceinline__ float fast_sqrt(float x) {
    return MCDX_DEV_OR_HOST(__builtin_amdgcn_sqrtf(x), sqrtf(x));
}
__host__ __device__ __forceinline__ float fast_exp2(float x) {
    return MCDX_DEV_OR_HOST(__builtin_amdgcn_exp2f(x), exp2f(x));
}

__host__ __device__ __forceinline__ double fast_rcp_f64(double x) {
    // v_rcp_f64 seed refined by one Newton step: relative error near 2^-52.
    const double y = MCDX_DEV_OR_HOST(__builtin_amdgcn_rcp(x), 1.0 / x);
    const double e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

__host__ __device__ __forceinline__ double fast_rsqrt_f64(double x) {
    // v_rsq_f64 seed refined by two Newton steps y <- y (3 - x y^2) / 2: relative error near
    // 2^-52 for positive normal x.
    double y = MCDX_DEV_OR_HOST(__builtin_amdgcn_rsq(x), 1.0 / __builtin_sqrt(x));
    const double hx = 0.5 * x;
    y = y * __builtin_fma(-hx * y, y, 1.5);
    y = y * __builtin_fma(-hx * y, y, 1.5);
    return y;
}

__host__ __device__ __forceinline__ bool is_tail(uint32_t bits) {
    return bits < kTailLo || bits > kTailHi;
}

// FP32-output Acklam on the raw Philox bits (Horner forms as normal.cpp:50-66, fused).
// The central rational is evaluated in FP64: near the tail break its denominator cancels
// from terms of size ~30 down to ~0.006, which FP32 Horner cannot carry (errors reach
// 1e-4). q = u - 0.5 is formed exactly from the integer bits, and FP64 FMAs issue at the
// FP32 rate on CDNA4, so the cost is a few conversions and a refined reciprocal. The
// well-conditioned tail stays in FP32. central_f32 is meaningful only for !is_tail(bits).
__host__ __device__ __forceinline__ float central_f32(uint32_t bits) {
    MCDX_ACKLAM_CENTRAL(double, MCDX_F64)
    const int32_t centred = (int32_t)(bits ^ 0x80000000u); // bits - 2^31, exact
    const double q = ((double)centred + 0.5) * 0x1p-32;     // exact
    const double r = q * q;
    double num = __builtin_fma(A1, r, A2);
    num = __builtin_fma(num, r, A3);
    num = __builtin_fma(num, r, A4);
    num = __builtin_fma(num, r, A5);
    num = __builtin_fma(num, r, A6);
    double den = __builtin_fma(B1, r, B2);
    den = __builtin_fma(den, r, B3);
    den = __builtin_fma(den, r, B4);
    den = __builtin_fma(den, r, B5);
    den = __builtin_fma(den, r, 1.0);
    return (float)(num * q * fast_rcp_f64(den));
}

__host__ __device__ __forceinline__ float tail_f32(uint32_t bits) {
    MCDX_ACKLAM_TAIL(float, MCDX_F32)
    const bool lower = bits < kTailLo;
    const uint32_t t = lower ? bits : ~bits; // ~bits = 2^32-1-bits: 1-u = (~bits+0.5)/2^32
    const float ut = __builtin_fmaf((float)t, 0x1p-32f, 0x1p-33f);
    // sqrt(-2 ln ut) = sqrt(-2 ln2 * log2 ut)
    const float z = fast_sqrt(-1.38629436111989061883f * fast_log2(ut));
    float num = __builtin_fmaf(C1, z, C2);
    num = __builtin_fmaf(num, z, C3);
    num = __builtin_fmaf(num, z, C4);
    num = __builtin_fmaf(num, z, C5);
    num = __builtin_fmaf(num, z, C6);
    float den = __builtin_fmaf(D1, z, D2);
    den = __builtin_fmaf(den, z, D3);
    den = __builtin_fmaf(den, z, D4);
    den = __builtin_fmaf(den, z, 1.0f);
    const float res = num * fast_rcp(den);
    return lower ? res : -res;
}

__host__ __device__ __forceinline__ float inv_normal_f32(uint32_t bits) {
    return is_tail(bits) ? tail_f32(bits) : central_f32(bits);
}

// B consecutive normal draws (draw0 .. draw0+B-1, stream 0) of one path. A wave64 holds a
// tail draw (probability 4.85% each) on almost every step, so evaluating the tail branch
// per step would cost the whole wave the tail's log/sqrt/rational nearly every time.
// Instead every draw takes the central branch, and each lane then works through only its
// own tail draws in the block: the wave loops max-over-lanes times (about 2.3 for B=8)
// instead of about 7.7.
template <int B>
__host__ __device__ __forceinline__ void normal_block_f32(uint32_t plo, uint32_t phh,
                                                          uint32_t draw0, uint32_t k0,
                                                          uint32_t k1, float (&z)[B]) {
    uint32_t bits[B];
    uint32_t pending = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        bits[j] = philox_w0(plo, phh, draw0 + (uint32_t)j, 0u, k0, k1);
        z[j] = central_f32(bits[j]);
        pending |= (uint32_t)is_tail(bits[j]) << j;
    }
    while (pending) {
        const int j = __builtin_ctz(pending);
        uint32_t b = bits[0];
#pragma unroll
        for (int jj = 1; jj < B; ++jj) b = (j == jj) ? bits[jj] : b;
        const float t = tail_f32(b);
#pragma unroll
        for (int jj = 0; jj < B; ++jj) z[jj] = (j == jj) ? t : z[jj];
        pending &= pending - 1;
    }
}

// FP64: the reference's arithmetic, operation for operation (normal.cpp:50-82).
__host__ __device__ __forceinline__ double uniform_open01_f64(uint32_t bits) {
    return ((double)bits + 0.5) / 4294967296.0;
}

__host__ __device__ __forceinline__ double normal_cdf_f64(double x) {
    return 0.5 * erfc(-x / 1.41421356237309504880168872420969808);
}

__host__ __device__ __forceinline__ double inv_normal_f64(double u) {
    MCDX_ACKLAM_CENTRAL(double, MCDX_F64)
    MCDX_ACKLAM_TAIL(double, MCDX_F64)
    constexpr double kLow = 0.02425;
    constexpr double kHigh = 1.0 - kLow;
    double z;
    if (u < kLow || kHigh < u) {
        const bool lower = u < kLow;
        const double t = sqrt(-2.0 * log(lower ? u : 1.0 - u));
        const double num = (((((C1 * t + C2) * t + C3) * t + C4) * t + C5) * t + C6);
        const double den = ((((D1 * t + D2) * t + D3) * t + D4) * t + 1.0);
        const double r = num / den;
        z = lower ? r : -r;
    } else {
        const double q = u - 0.5;
        const double r = q * q;
        const double num = (((((A1 * r + A2) * r + A3) * r + A4) * r + A5) * r + A6) * q;
        const double den = (((((B1 * r + B2) * r + B3) * r + B4) * r + B5) * r + 1.0);
        z = num / den;
    }
    const double error = normal_cdf_f64(z) - u;
    const double sqrt_2pi = 0x1.40d931ff62705p+1; // sqrt(2.0 * pi) as the FP64 sqrt rounds it
    const double r = error * sqrt_2pi * exp(0.5 * z * z);
    z -= r / (1.0 + 0.5 * z * r);
    return z;
}

// ---------------------------------------------------------------------------
// Welford accumulator (FP64) and the Chan merge, stats.hpp:23-65
// ---------------------------------------------------------------------------
struct Welford {
    uint64_t n;
    double mean;
    double m2;
};

__host__ __device__ __forceinline__ Welford welford_merge(Welford a, Welford b) {
    if (b.n == 0) return a;
    if (a.n == 0) return b;
    const uint64_t nc = a.n + b.n;
    const double delta = b.mean - a.mean;
    const double w = (double)b.n / (double)nc;
    Welford r;
    r.n = nc;
    r.mean = a.mean + delta * w;
    r.m2 = a.m2 + b.m2 + delta * delta * (double)a.n * (double)b.n / (double)nc;
    return r;
}

// Canonical work decomposition shared by every MC kernel so that a contract's result is
// independent of grid size and batching: a tile is 1024 consecutive path indices handled
// by 256 lanes x 4 paths (path = tile*1024 + k*256 + lane); each lane folds its paths in k
// order, the 256 lane accumulators merge along a pairwise tree in lane order, and tiles
// merge by the two-level scheme in final_merge (reduce.hip).
constexpr int kTileLanes = 256;
constexpr int kPathsPerLane = 4;
constexpr int kTilePaths = kTileLanes * kPathsPerLane;
constexpr int kMergeGroups = 1024;

#if defined(__HIPCC__)
// Pairwise tree over the 256 lanes of a block: within each wave by lane offset 1..32, then
// across the four waves (w0+w1, w2+w3, then both). Returns the tile total on thread 0.
__device__ __forceinline__ Welford block_merge_256(Welford w, Welford* lds4) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Welford o;
        o.n = __shfl_down(w.n, off, 64);
        o.mean = __shfl_down(w.mean, off, 64);
        o.m2 = __shfl_down(w.m2, off, 64);
        if ((lane & (2 * off - 1)) == 0) w = welford_merge(w, o);
    }
    const int wave = threadIdx.x >> 6;
    if (lane == 0) lds4[wave] = w;
    __syncthreads();
    Welford r = {0, 0.0, 0.0};
    if (threadIdx.x == 0) r = welford_merge(welford_merge(lds4[0], lds4[1]),
                                            welford_merge(lds4[2], lds4[3]));
    return r;
}

// The same tree, with the tile total returned to every lane and no lane-divergent branch:
// each wave broadcasts its lane-0 total with readfirstlane, every lane of the wave writes
// that same value to the wave's LDS slot, and every thread folds the four slots. Used in
// loops that run the merge more than once, where a thread-0-only region next to the
// barriers lets the compiler split the loop so lanes would re-run iterations out of step.
__device__ __forceinline__ uint64_t bcast_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double bcast_f64(double v) {
    return __builtin_bit_cast(double, bcast_u64(__builtin_bit_cast(uint64_t, v)));
}

__device__ __forceinline__ Welford block_merge_256_all(Welford w, Welford* lds4) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Welford o;
        o.n = __shfl_down(w.n, off, 64);
        o.mean = __shfl_down(w.mean, off, 64);
        o.m2 = __shfl_down(w.m2, off, 64);
        const Welford m = welford_merge(w, o);
        const bool take = (lane & (2 * off - 1)) == 0;
        w.n = take ? m.n : w.n;
        w.mean = take ? m.mean : w.mean;
        w.m2 = take ? m.m2 : w.m2;
    }
    Welford wave_total;
    wave_total.n = bcast_u64(w.n);
    wave_total.mean = bcast_f64(w.mean);
    wave_total.m2 = bcast_f64(w.m2);
    lds4[threadIdx.x >> 6] = wave_total;
    __syncthreads();
    const Welford r = welford_merge(welford_merge(lds4[0], lds4[1]), welford_merge(lds4[2], lds4[3]));
    __syncthreads(); // the slots may be rewritten by the caller's next merge
    return r;
}
#endif

} // namespace mcdx
