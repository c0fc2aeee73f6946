// Device primitives shared by every path kernel (gfx950, wave64).
//
//  * Philox4x32-10 in registers, one counter per lane: key {seed_lo, seed_hi},
//    counter {path_lo, path_hi, draw, stream}, output word 0 only
//    (reference src/core/rng.cpp:20-74). The key schedule and every counter word
//    that is uniform across the wave (draw, stream) fold onto the scalar unit.
//  * Inverse normal CDF, two precisions:
//      - FP64: Acklam + one Halley step against 0.5*erfc(-x/sqrt2), the reference's
//        exact arithmetic (src/core/normal.cpp:50-82), built without FP contraction.
//      - FP32: this engine's own polynomials in y = log2(4m(1-m)), m = min(u, 1-u)
//        (see inv_normal_f32 below): no division, one v_log_f32, a tail branch for 0.097%
//        of draws decided on the integer bits, and m formed from min(bits, ~bits) so 1-u
//        is exact (no u rounding to 1.0f). At most 1.89e-7 of (1 + |z|) from the FP64
//        result, measured over all 2^32 words.
//  * FP64 Welford accumulators and the Chan-Golub-LeVeque merge
//    (include/mcd/core/stats.hpp:21-65), merged along a fixed pairwise tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "acklam.hpp"

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
        // Round 1's left word comes from c1, c2 and the key only: plain XORs, so that when
        // those are wave-uniform (a path high word shared by the wave, the draw index, the
        // seed) it and the next round's M0 product stay on the scalar unit.
        const uint32_t n0 = r == 0 ? (uint32_t)(p1 >> 32) ^ c1 ^ k0 : xor3((uint32_t)(p1 >> 32), c1, k0);
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
#define MCDX_F32(x) ((float)(x))
#define MCDX_F64(x) ((double)(x))

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

// ---- FP32 inverse normal -------------------------------------------------------------------
// Phi^-1(u), u = (bits + 0.5) / 2^32, for the FP32 generation mode: an approximation of this
// library (the FP64 mode keeps the reference's Acklam + Halley). With m = min(u, 1 - u) and
// y = log2(4 m (1 - m)) = log2(1 - 4 q^2), q = u - 1/2:
//   z = q * P(y)                        when m >= m0 = 4.83e-4 (99.9% of draws),
//   z = sign(q) * T(sqrt(-y) - kTailC)  otherwise,
// with P (degree 9) and T (degree 8) weighted minimax polynomials fitted in FP64 against
// sqrt(2) erfinv(2u - 1) (relative error 7.7e-8 and 6.4e-9; end to end in FP32 arithmetic at
// most 1.89e-7 of (1 + |z|) against the FP64 Acklam + Halley, measured over all 2^32 words). The change of variable removes the rational's division and moves
// the branch point from Acklam's u = 0.02425 (4.85% of draws) to m0 (0.097%), so a wave
// rarely takes the tail branch.
constexpr uint32_t kTailInt = 2073811u; // tail iff min(bits, ~bits) < kTailInt, i.e. m < m0
constexpr float kTailC = 4.285285692405615f;

// min(bits, ~bits): (m * 2^32 - 0.5) for m = min(u, 1 - u), since 1 - u = (~bits + 0.5) / 2^32.
__host__ __device__ __forceinline__ uint32_t fold_half(uint32_t bits) {
    return bits ^ (uint32_t)((int32_t)bits >> 31);
}

__host__ __device__ __forceinline__ bool is_tail(uint32_t bits) { return fold_half(bits) < kTailInt; }

// y = log2(4 m (1 - m)) from the folded bits (4m formed exactly scaled, one rounding each).
__host__ __device__ __forceinline__ float half_log2(float mf) {
    const float m4 = __builtin_fmaf(mf, 0x1p-30f, 0x1p-31f);
    return fast_log2(__builtin_fmaf(-0.25f * m4, m4, m4));
}

// z carries the sign of q = u - 1/2: za XOR (sign bit of ~bits).
__host__ __device__ __forceinline__ float signed_by_bits(float za, uint32_t bits) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, za) ^ (~bits & 0x80000000u));
}

// P(y): z = |q| * P(y) in the central range.
__host__ __device__ __forceinline__ float central_poly(float y) {
    float p = __builtin_fmaf(5.28648458342218e-10f, y, 2.0134757861001162e-08f);
    p = __builtin_fmaf(p, y, 2.446661255817162e-07f);
    p = __builtin_fmaf(p, y, -1.4628129463289952e-07f);
    p = __builtin_fmaf(p, y, -2.6267098292009905e-05f);
    p = __builtin_fmaf(p, y, -9.876400872599334e-05f);
    p = __builtin_fmaf(p, y, 0.002176771406084299f);
    p = __builtin_fmaf(p, y, 0.015678392723202705f);
    p = __builtin_fmaf(p, y, -0.45487093925476074f);
    return __builtin_fmaf(p, y, 2.5066280364990234f);
}

// Meaningful only for !is_tail(bits); mi = fold_half(bits).
__host__ __device__ __forceinline__ float central_f32(uint32_t bits, uint32_t mi) {
    const float mf = (float)mi;
    return signed_by_bits(__builtin_fmaf(mf, -0x1p-32f, 0.5f) * central_poly(half_log2(mf)), bits);
}

// T: |z| in the tail range, from y.
__host__ __device__ __forceinline__ float tail_poly(float y) {
    const float t = fast_sqrt(-y) - kTailC;
    float p = __builtin_fmaf(-1.7373583887092536e-06f, t, 1.4796261211813544e-06f);
    p = __builtin_fmaf(p, t, 1.7690343156573363e-05f);
    p = __builtin_fmaf(p, t, -9.297743963543326e-05f);
    p = __builtin_fmaf(p, t, 0.0003736739745363593f);
    p = __builtin_fmaf(p, t, -0.001273712026886642f);
    p = __builtin_fmaf(p, t, 0.0027184723876416683f);
    p = __builtin_fmaf(p, t, 1.186617374420166f);
    return __builtin_fmaf(p, t, 4.813599109649658f);
}

__host__ __device__ __forceinline__ float tail_f32(uint32_t bits) {
    return signed_by_bits(tail_poly(half_log2((float)fold_half(bits))), bits);
}

// Central branch always; the tail polynomial only in lanes that need it (0.097% of draws).
__host__ __device__ __forceinline__ float inv_normal_f32(uint32_t bits) {
    const uint32_t mi = fold_half(bits);
    float z = central_f32(bits, mi);
    if (mi < kTailInt) z = tail_f32(bits);
    return z;
}

// B consecutive normal draws (draw0 .. draw0+B-1, stream 0) of one path. Every draw takes
// the central branch and the block keeps only the smallest folded value; a lane re-examines
// its draws for tails only when that minimum is in the tail range (0.8% of lane-blocks for
// B = 8 at 0.097% per draw), so the common path pays one min per draw for the tail test.
template <int B>
__host__ __device__ __forceinline__ void normal_block_f32(uint32_t plo, uint32_t phh,
                                                          uint32_t draw0, uint32_t k0,
                                                          uint32_t k1, float (&z)[B]) {
    uint32_t bits[B];
    uint32_t lowest = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        bits[j] = philox_w0(plo, phh, draw0 + (uint32_t)j, 0u, k0, k1);
        const uint32_t mi = fold_half(bits[j]);
        z[j] = central_f32(bits[j], mi);
        lowest = lowest < mi ? lowest : mi;
    }
    if (lowest < kTailInt) {
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (is_tail(bits[j])) z[j] = tail_f32(bits[j]);
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
// Fixed-order merge of n_tiles consecutive tile partials by one block of kMergeGroups threads:
// group g folds the tiles chunk_bounds(g, n_tiles, kMergeGroups) left to right (the
// reference's static partition, thread_pool.cpp:9-18), then a pairwise tree over the groups
// in index order. The total is valid on thread 0.
__device__ __forceinline__ Welford merge_tiles_fixed_order(const Welford* __restrict__ partials,
                                                           uint64_t n_tiles, Welford* lds) {
    const uint32_t g = threadIdx.x;
    const uint64_t q = n_tiles / kMergeGroups, rem = n_tiles % kMergeGroups;
    const uint64_t b = g * q + (g < rem ? g : rem);
    const uint64_t e = b + q + (g < rem ? 1 : 0);
    Welford acc = {0, 0.0, 0.0};
    for (uint64_t t = b; t < e; ++t) acc = welford_merge(acc, partials[t]);
    lds[g] = acc;
    __syncthreads();
    for (uint32_t s = 1; s < kMergeGroups; s <<= 1) {
        if ((g & (2 * s - 1)) == 0) lds[g] = welford_merge(lds[g], lds[g + s]);
        __syncthreads();
    }
    return lds[0];
}

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
