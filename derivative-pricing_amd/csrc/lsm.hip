// Longstaff-Schwartz American pricing: HBM path store + per-date device regression.
//
// Reference behaviour restated here (src/pricers/lsm.cpp):
//   simulate_paths           :35-58  spot at every date, draw = step, GBM step per date
//   backward induction       :83-123 cf *= exp(-r dt); regress the ITM paths' discounted
//                                     cash flows on 4 weighted Laguerre terms of S/K; skip the
//                                     date when fewer than 4 paths are ITM; exercise when the
//                                     intrinsic value beats the fitted continuation
//   settlement               :125-148 cf *= disc once more; Welford; inception rule
//   frozen-policy replay     :151-213
//   laguerre_basis           :21-25, householder_least_squares src/core/linalg.cpp:7-71
//
// Layout: the store is SoA [date][path] (dates 1..M; the t=0 column is the constant S0),
// FP32 in FP32 mode (2^20 paths x 50 dates = 210 MB), FP64 in FP64 mode. Each path's cash
// flow is kept discounted to t = 0, intrinsic(S at its exercise date e) x disc^e with the
// powers of the one-date discount factor formed by repeated multiplication (the reference's
// cf *= disc per date, lsm.cpp:84-86, up to its rounding), so a date's kernel writes a path's
// value only when it exercises; a date's regression target is that value over disc^date, a
// factor common to the date's rows that scales the coefficients once.
//
// Regression: a date's rows [basis(S/K)] are those of the paths in the money at that date,
// fixed by the store alone; only the target depends on the later dates' decisions. So every
// date's preconditioner P (upper triangular, P^T P close to the rows' Gram matrix) is formed
// before the backward induction, for all dates at once, by a Cholesky-QR step over a path
// sample. Each date of the backward induction then solves the later date's regression,
// applies its exercise decision, and accumulates the normal equations of the preconditioned
// rows w = P^-T basis (G' = sum w w^T, close to the identity, and q = sum w cf), reduced per
// block: one launch runs every date when the run's blocks fit on the device at once (they
// exchange each date's sums through device-scope memory), else one launch per date, with the
// same bits. The coefficients are P^-1 G'^-1 q with G' by Cholesky: the
// least-squares solution of the reference's Householder sweep (linalg.cpp:7-71) to within
// rounding of order the rows' condition number (about 2e6 at the first dates of a 50-date
// put) times 1e-16, since the preconditioned problem is well conditioned.

#include <math.h>

#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

#include "mlp.hpp"
#include "qmc_store.hpp"
#include "runtime.hpp"

namespace mcdx {

constexpr int kLsmBasis = 4;
constexpr int kLsmPathLanes = 256; // one path per lane: store and cash-flow init kernels

// A date's preconditioner: P upper triangular (row-major packed 10 values) with
// P^T P close to the Gram matrix of the date's in-the-money rows, and the reciprocals of its
// diagonal (0 where the diagonal is 0: that column is dropped, as linalg.cpp:65-68 drops a
// zero-diagonal column in the back-substitution).
struct DateFactor {
    double r[10];
    double inv_d[4];
};

__host__ __device__ __forceinline__ int ridx(int i, int j) { // i <= j
    return i * kLsmBasis - (i * (i - 1)) / 2 + (j - i);
}

// Sums per date: the Gram matrix of the rows (upper triangle, packed like P) and the row
// count (preconditioner pass); G' = W^T W (packed), q = W^T y and the row count (backward
// induction).
constexpr int kGram = 11;
constexpr int kNe = 15;
constexpr int kLsmNeStride = 256; // block sums per value in the ne buffer (at most 256 blocks)

// The reference's basis; x^3 / 6 is formed as x^3 * (1/6) (within an ulp of the division,
// without its dependent FP64 division sequence). Every kernel that fits or replays a policy
// evaluates this same function, so fits and replays agree bit for bit.
// The basis at x given its weight e = e^{-x/2}.
__host__ __device__ __forceinline__ void laguerre4_weighted(double x, double e, double* out) {
    constexpr double kSixth = 1.0 / 6.0;
    const double x2 = x * x;
    out[0] = e;
    out[1] = e * (1.0 - x);
    out[2] = e * (1.0 - 2.0 * x + 0.5 * x2);
    out[3] = e * (1.0 - 3.0 * x + 1.5 * x2 - x2 * x * kSixth);
}
__host__ __device__ __forceinline__ void laguerre4(double x, double* out) { // lsm.cpp:21-25
    laguerre4_weighted(x, exp(-x / 2.0), out);
}

__host__ __device__ __forceinline__ double intrinsic_v(double s, double k, int type) {
    const double phi = type == 0 ? 1.0 : -1.0;
    const double v = phi * (s - k);
    return (v < 0.0) ? 0.0 : v;
}

struct LsmArgs {
    double spot, strike, disc, dt;
    double disc_m;         // disc^M by M multiplications (disc_pow)
    double inv_strike;     // 1/K: the regression basis is taken at S * (1/K) (one multiply
                           // per use instead of a division; every decision uses the same)
    double drift, vsd;     // FP64: per-step GBM constants (gbm.cpp:8-10 with dt)
    float mu2, sig2;       // FP32: log2-space step
    uint64_t seed, n_paths;
    int32_t m;             // exercise dates (monitoring points)
    uint32_t blocks;       // backward-induction layout: blocks and paths per lane
    uint32_t per_lane;
    uint64_t sample_n;     // the first sample_n paths precondition every date
    uint8_t type;
    void* store;           // [date 1..M][path], float or double
    double* val;           // [path]: the path's cash flow discounted to t = 0
    double* dpow;          // [M + 1]: disc^k by repeated multiplication
    double* coeffs;        // [(M+1) x 4] device
    uint8_t* has;          // [M+1] device
    double* gram;          // [date 1..M-1][kGram][blocks]: block Gram sums of each date's rows
    DateFactor* factors;   // [M+1]: each date's preconditioner (dates 1..M-1)
    double* ne;            // [2][kNe][kLsmMaxBlocks]: block sums of the normal equations of
                           // even and odd dates (zero past the blocks)
    Welford* partials;     // [tiles]
    // Optional capture of the regression samples for the network, [date 1..M-1][path]:
    // features (4 (S/K - 1), 2 date/M - 1) and target 10 cf/K (NaN when out of the money).
    float* cap_f1;
    float* cap_f2;
    float* cap_y;
    // A batch of independent runs of the same contract on different stores (grid y = run,
    // grid z for the Gram kernel): the distance between consecutive runs' store (bytes),
    // exercise states (paths), Gram sums, normal-equation sums and tile partials; coeffs, has
    // and factors advance by (M+1) x 4, M+1 and M+1. Zero for a single run.
    uint64_t run_store_bytes, run_cf, run_gram, run_ne, run_tiles;
    // The one-launch backward induction's date counters of each run: block b adds to counter
    // b mod kLsmSyncGroups (kLsmSyncWords words apart, so blocks contend in small groups);
    // zeroed by the preconditioner merge that precedes it.
    uint32_t* sync;
};
constexpr uint32_t kLsmSyncGroups = 32;
constexpr uint32_t kLsmSyncWords = 64; // one counter per 256 bytes
constexpr uint32_t kLsmSyncStride = kLsmSyncGroups * kLsmSyncWords; // words per run

// The arguments of run `y` of a batch (itself for a single run). The capture buffers are not
// batched: only single runs capture.
__device__ __forceinline__ void select_run(LsmArgs& a, uint64_t y) {
    a.store = (char*)a.store + y * a.run_store_bytes;
    a.val += y * a.run_cf;
    a.coeffs += y * (uint64_t)(a.m + 1) * 4;
    a.has += y * (uint64_t)(a.m + 1);
    a.gram += y * a.run_gram;
    a.factors += (a.run_gram ? y : 0) * (uint64_t)(a.m + 1);
    a.ne += y * a.run_ne;
    a.partials += y * a.run_tiles;
    a.sync += y * kLsmSyncStride;
}
__device__ __forceinline__ void select_run(LsmArgs& a) { select_run(a, blockIdx.y); }

constexpr double kFeatScale = 4.0;  // f1 = kFeatScale * (S/K - 1)
constexpr double kTargetScale = 10.0; // y = kTargetScale * cf / K

template <bool FP64>
__device__ __forceinline__ double store_get(const LsmArgs& a, int date, uint64_t p) {
    const uint64_t idx = (uint64_t)(date - 1) * a.n_paths + p;
    if constexpr (FP64) return ((const double*)a.store)[idx];
    else return (double)((const float*)a.store)[idx];
}


// Forward simulation into the store (lsm.cpp:35-58), one path per lane; coalesced
// [date][path] writes. FP32 mode steps in log2 space like the path kernels.
template <bool FP64>
__global__ __launch_bounds__(256) void lsm_paths_kernel(LsmArgs a) {
    // Launched with 256 lanes per block: the block's paths share p >> 32, a scalar here.
    const uint64_t base = (uint64_t)blockIdx.x * 256u;
    const uint64_t p = base + threadIdx.x;
    if (p >= a.n_paths) return;
    const uint32_t plo = (uint32_t)p, phh = (uint32_t)(base >> 32);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    if constexpr (FP64) {
        double s = a.spot;
        double* st = (double*)a.store;
        for (int step = 0; step < a.m; ++step) {
            const double z = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
            s = s * exp(a.drift + a.vsd * z);
            st[(uint64_t)step * a.n_paths + p] = s;
        }
    } else {
        float x = 0.f;
        float* st = (float*)a.store;
        const float s0 = (float)a.spot;
        int step = 0;
        constexpr int B = 8;
        for (; step + B <= a.m; step += B) {
            float z[B];
            normal_block_f32<B>(plo, phh, (uint32_t)step, k0, k1, z);
#pragma unroll
            for (int j = 0; j < B; ++j) {
                x = x + __builtin_fmaf(a.sig2, z[j], a.mu2);
                st[(uint64_t)(step + j) * a.n_paths + p] = s0 * fast_exp2(x);
            }
        }
        for (; step < a.m; ++step) {
            const float z = inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1));
            x = x + __builtin_fmaf(a.sig2, z, a.mu2);
            st[(uint64_t)step * a.n_paths + p] = s0 * fast_exp2(x);
        }
    }
}

// disc^k by k multiplications from 1: the discount powers backward induction applies (the
// table dpow holds the same values).
__host__ __device__ __forceinline__ double disc_pow(double disc, int k) {
    double d = 1.0;
    for (int i = 0; i < k; ++i) d *= disc;
    return d;
}

// Every path's value at maturity, intrinsic x disc^M (lsm.cpp:76-81, discounted as :84-86 and
// :125-127 will); block 0 of each run also zeroes the run's normal-equation sums and, when
// fitting, clears the run's policy (no coefficients at any date yet), and of run 0 forms the
// discount powers.
template <bool FP64>
__global__ __launch_bounds__(256) void lsm_init_cf_kernel(LsmArgs a, int fit) {
    select_run(a);
    if (blockIdx.x == 0) {
        for (int i = (int)threadIdx.x; i <= a.m && fit; i += (int)blockDim.x) {
            a.has[i] = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) a.coeffs[4 * i + j] = 0.0;
        }
        for (int i = (int)threadIdx.x; i < 2 * kNe * kLsmNeStride; i += (int)blockDim.x) a.ne[i] = 0.0;
        if (threadIdx.x == 0 && blockIdx.y == 0) {
            double d = 1.0;
            for (int k = 0; k <= a.m; ++k) {
                a.dpow[k] = d;
                d *= a.disc;
            }
        }
    }
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    a.val[p] = intrinsic_v(store_get<FP64>(a, a.m, p), a.strike, a.type) * a.disc_m;
}

// ---- Backward induction -------------------------------------------------------------------
// Path layout: B = min(256, ceil(n / 4096)) blocks of 512 lanes (two waves per SIMD); lane t
// of block b owns paths ((j * B + b) * 512 + t) for j < P = ceil(n / (512 B)), coalesced per j
// (P = 8 at 2^20 paths). Before the dates, the Gram kernels precondition every date at once
// (grid: blocks x dates) from the rows of the first min(n, 2^16) paths; then, per date (in
// one resident launch, or one launch each): the later date's coefficients from the B blocks'
// normal-equation sums, the decision at the later date, discount, and the normal equations of
// this date's preconditioned rows, summed over the block. The layout depends on n only and
// every sum runs in a fixed order, so the coefficients are bitwise reproducible.
constexpr int kLsmLanes = 512;
constexpr int kLsmWaves = kLsmLanes / 64;
constexpr uint32_t kLsmMaxBlocks = kLsmNeStride;
static_assert(kLsmMaxBlocks <= 256 && kLsmMaxBlocks <= (uint32_t)kLsmLanes,
              "the merges load at most four blocks' sums per lane of wave 0, one per lane otherwise");
constexpr uint32_t kLsmChunk = 8; // paths per lane whose loads are issued together
// The sample that preconditions each date: the first kLsmSample paths (independent paths, so
// a random sample of the date's rows; at 2^16 paths its Gram matrix is within about 1% of
// the full one's in every direction, so the preconditioned rows stay within about 1% of
// orthonormal). A date whose sample holds fewer than kLsmMinSample in-the-money rows is
// preconditioned from every path instead.
constexpr uint64_t kLsmSample = 1ull << 16;
constexpr double kLsmMinSample = 1024.0;
// Diagonal shift of the preconditioner's Cholesky, relative to the Gram trace: it keeps the
// factorisation defined for rank-deficient or extremely ill-conditioned rows and is far
// below the smallest eigenvalue ratio of the Laguerre rows (about 2e-13 at the first dates
// of a 50-date put).
constexpr double kLsmShift = 0x1p-46;

__host__ __device__ __forceinline__ uint64_t lane_path(uint32_t j, uint32_t block,
                                                       uint32_t blocks, uint32_t t) {
    return ((uint64_t)j * blocks + block) * kLsmLanes + t;
}

// Exercise decision at a later date with coefficients beta (lsm.cpp:110-122): whether the
// path exercises.
__device__ __forceinline__ bool decide(const LsmArgs& a, double s1, const double* beta) {
    const double ex = intrinsic_v(s1, a.strike, a.type);
    if (ex > 0.0) {
        double b[4];
        laguerre4(s1 * a.inv_strike, b);
        double cont = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) cont += beta[j] * b[j];
        return ex > cont;
    }
    return false;
}

// A date's continuation value without its weight, sum_j beta_j L_j(x), as the cubic
// c0 + c1 x + c2 x^2 + c3 x^3, and the bound sum_j |beta_j| (1 + x)^j on sum_j |beta_j L_j(x)|
// and on the cubic's terms sum_k |c_k| x^k (for x >= 0, |L_j(x)| <= (1 + x)^j term by term).
struct ContPoly {
    double c[4], b[4];
};
__device__ __forceinline__ ContPoly cont_poly(const double* beta) {
    constexpr double kSixth = 1.0 / 6.0;
    ContPoly p;
    p.c[0] = ((beta[0] + beta[1]) + beta[2]) + beta[3];
    p.c[1] = -((beta[1] + 2.0 * beta[2]) + 3.0 * beta[3]);
    p.c[2] = 0.5 * beta[2] + 1.5 * beta[3];
    p.c[3] = -beta[3] * kSixth;
#pragma unroll
    for (int j = 0; j < 4; ++j) p.b[j] = fabs(beta[j]);
    return p;
}

// decide() with the same result, cheaper: the continuation value is estimated as the weight
// e^{-x/2} from the hardware FP32 exponential (relative error below 1.2e-7 + 9e-8 x: one ulp of
// v_exp_f32 and the FP32 argument's rounding) times the cubic in FP64. The estimate differs
// from decide()'s FP64 value by at most that relative error of the estimate plus the roundings
// of both evaluations and of the cubic's coefficients, each a few 1e-16 of the bound
// e sum_j |beta_j| (1 + x)^j (the sum cancels: its terms are 40-1000 times its value). Where
// the intrinsic value is farther from the estimate than 4e-6 (1 + x) |estimate| + 1e-12
// bound, the estimate's side of the boundary is decide()'s; the remaining paths (within a few
// 1e-6 of the boundary, or with x outside [0, 64]) are decided by decide() itself.
// Without branches: whether the path exercises by the estimate (*out) and whether the
// estimate cannot decide it (returned: the caller then calls decide()).
__device__ __forceinline__ bool decide_screen(const LsmArgs& a, double s1, const ContPoly& cp,
                                              bool* out) {
    const double ex = intrinsic_v(s1, a.strike, a.type);
    const double x = s1 * a.inv_strike;
    const double e = (double)fast_exp2((float)x * -0.721347520444481703680f); // e^{-x/2}
    const double est = e * __builtin_fma(__builtin_fma(__builtin_fma(cp.c[3], x, cp.c[2]), x, cp.c[1]), x, cp.c[0]);
    const double x1 = 1.0 + x;
    const double bound = e * __builtin_fma(__builtin_fma(__builtin_fma(cp.b[3], x1, cp.b[2]), x1, cp.b[1]), x1, cp.b[0]);
    const bool sure = x >= 0.0 && x <= 64.0 && fabs(ex - est) > 4e-6 * x1 * fabs(est) + 1e-12 * bound;
    const bool itm = ex > 0.0;
    *out = itm && ex > est;
    return itm && !sure;
}

// A row for the preconditioner only: the basis with its weight e^{-x/2} from the hardware FP32
// exponential, which scales each row by 1 + O(1e-7); the preconditioned rows of the exact
// basis are then orthonormal to within that.
__device__ __forceinline__ bool design_row_approx(const LsmArgs& a, bool in, double s0, double* row) {
    constexpr double kSixth = 1.0 / 6.0;
    const bool itm = in && intrinsic_v(s0, a.strike, a.type) > 0.0;
    const double x = s0 * a.inv_strike;
    const double e = itm ? (double)fast_exp2((float)x * -0.721347520444481703680f) : 0.0;
    const double x2 = x * x;
    row[0] = e;
    row[1] = e * (1.0 - x);
    row[2] = e * (1.0 - 2.0 * x + 0.5 * x2);
    row[3] = e * (1.0 - 3.0 * x + 1.5 * x2 - x2 * x * kSixth);
    return itm;
}

// One path's regression row (lsm.cpp:88-100): the Laguerre basis of S/K when the path is in
// range and in the money, else zeros; returns whether it is.
__device__ __forceinline__ bool design_row(const LsmArgs& a, bool in, double s0, double* row) {
    const bool itm = in && intrinsic_v(s0, a.strike, a.type) > 0.0;
    double b[4];
    laguerre4(s0 * a.inv_strike, b);
#pragma unroll
    for (int k = 0; k < 4; ++k) row[k] = itm ? b[k] : 0.0;
    return itm;
}

// The network's training sample of a path in range at `date`: features (4 (S/K - 1),
// 2 date/M - 1) and target 10 cf/K (NaN when out of the money).
__device__ __forceinline__ void capture_row(const LsmArgs& a, int date, uint64_t p, bool itm,
                                            double s0, double cf) {
    const uint64_t c = (uint64_t)(date - 1) * a.n_paths + p;
    a.cap_f1[c] = (float)(kFeatScale * (s0 / a.strike - 1.0));
    a.cap_f2[c] = (float)(2.0 * (double)date / (double)a.m - 1.0);
    a.cap_y[c] = itm ? (float)(kTargetScale * cf / a.strike) : __builtin_nanf("");
}

// DPP lane permutation of a double (both halves moved by the same control).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sums of NV doubles over the 64 lanes of a wave, by transposition through the wave's own LDS
// region wl ([NV][65], padded against bank conflicts): every lane writes its values, lane
// 4 i + q (i < NV) adds value i of lanes 16 q .. 16 q + 15 in lane order, and two quad
// exchanges add the quarters as (q0 + q1) + (q2 + q3) (both lanes of a pair add the same two
// operands). Lane 4 i returns the wave's total of value i. The region is the wave's alone, so
// a wave-level barrier (no block barrier) orders the writes before the reads; a later call
// rewrites it only after this wave has read it.
constexpr int kLsmWaveStride = 65;
template <int NV>
__device__ __forceinline__ double wave_sums(const double (&v)[NV], double* wl) {
    const int l = (int)(threadIdx.x & 63);
#pragma unroll
    for (int i = 0; i < NV; ++i) wl[i * kLsmWaveStride + l] = v[i];
    // the other lanes' writes visible to this wave's reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int i = l >> 2, q = l & 3;
    double s = 0.0;
    if (i < NV) {
        const double* r = wl + i * kLsmWaveStride + q * 16;
        s = r[0];
#pragma unroll
        for (int k = 1; k < 16; ++k) s += r[k];
    }
    s += dpp_f64<0xB1>(s); // quad_perm [1, 0, 3, 2]
    s += dpp_f64<0x4E>(s); // quad_perm [2, 3, 0, 1]
    return s;
}

// Block-wide sums of NV doubles: wave sums, the wave totals through LDS (one barrier), then
// lane i < NV of every wave adds value i's kLsmWaves totals by a fixed pairwise tree and
// returns it. The wave regions are rewritten only by their own wave; the totals array is
// rewritten by a later call only after a barrier that follows every read of it.
constexpr int kLsmWaveRegion = kNe * kLsmWaveStride;
constexpr int kLsmLdsTotals = kLsmWaves * kLsmWaveRegion;      // [wave][kNe] totals
constexpr int kLsmLdsRes = kLsmLdsTotals + kLsmWaves * kNe;     // the solve's [kNe + 5]
constexpr int kLsmLdsDoubles = kLsmLdsRes + kNe + 5;
template <int NV>
__device__ __forceinline__ double block_sums(const double (&v)[NV], double* lds) {
    static_assert(NV <= kNe && (kLsmWaves & (kLsmWaves - 1)) == 0, "exchange layout");
    const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const double ws = wave_sums<NV>(v, lds + w * kLsmWaveRegion);
    double* tot = lds + kLsmLdsTotals;
    if ((l & 3) == 0 && (l >> 2) < NV) tot[w * NV + (l >> 2)] = ws;
    __syncthreads();
    double t[kLsmWaves];
#pragma unroll
    for (int k = 0; k < kLsmWaves; ++k) t[k] = l < NV ? tot[k * NV + l] : 0.0;
#pragma unroll
    for (int h = kLsmWaves / 2; h >= 1; h /= 2) {
#pragma unroll
        for (int k = 0; k < h; ++k) t[k] = t[2 * k] + t[2 * k + 1];
    }
    return t[0];
}

// The preconditioned row w = P^-T row by forward substitution; a dropped column (zero
// diagonal) gives w_i = 0.
__device__ __forceinline__ void precondition_row(const DateFactor& f, const double* row, double* w) {
    w[0] = row[0] * f.inv_d[0];
    w[1] = __builtin_fma(-f.r[ridx(0, 1)], w[0], row[1]) * f.inv_d[1];
    w[2] = __builtin_fma(-f.r[ridx(1, 2)], w[1], __builtin_fma(-f.r[ridx(0, 2)], w[0], row[2])) *
           f.inv_d[2];
    w[3] = __builtin_fma(-f.r[ridx(2, 3)], w[2],
                         __builtin_fma(-f.r[ridx(1, 3)], w[1],
                                       __builtin_fma(-f.r[ridx(0, 3)], w[0], row[3]))) *
           f.inv_d[3];
}

// ---- Preconditioners of every date (before the backward induction) ----------------------
// Gram sums of the rows of one date over paths p < lim, as (j gb + b) kLsmLanes + t for
// j < per (block b of gb, lane t), added to g.
template <bool FP64>
__device__ __forceinline__ void gram_rows(const LsmArgs& a, int date, uint32_t gb, uint32_t b,
                                          uint32_t per, uint64_t lim, double (&g)[kGram]) {
    for (uint32_t j0 = 0; j0 < per; j0 += kLsmChunk) {
        double s0[kLsmChunk];
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = ((uint64_t)(j0 + u) * gb + b) * kLsmLanes + threadIdx.x;
            s0[u] = store_get<FP64>(a, date, (j0 + u < per && p < lim) ? p : 0);
        }
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = ((uint64_t)(j0 + u) * gb + b) * kLsmLanes + threadIdx.x;
            double row[4];
            const bool itm = design_row_approx(a, j0 + u < per && p < lim, s0[u], row);
#pragma unroll
            for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
                for (int k = i; k < kLsmBasis; ++k) g[ridx(i, k)] = __builtin_fma(row[i], row[k], g[ridx(i, k)]);
            }
            g[10] += itm ? 1.0 : 0.0;
        }
    }
}

// Grid (gb, M - 1, runs): block b of date y + 1 sums the Gram matrix and the count of the
// rows of the sample (the first min(n, 2^16) paths). gb <= kLsmGramBlocks.
constexpr uint32_t kLsmGramBlocks = 32;
template <bool FP64>
__global__ __launch_bounds__(kLsmLanes) void lsm_gram_kernel(LsmArgs a, uint32_t gb, uint32_t per) {
    select_run(a, blockIdx.z);
    const int date = (int)blockIdx.y + 1;
    __shared__ double lds[kLsmLdsDoubles];
    double g[kGram];
#pragma unroll
    for (int k = 0; k < kGram; ++k) g[k] = 0.0;
    gram_rows<FP64>(a, date, gb, blockIdx.x, per, a.sample_n, g);
    const double t = block_sums<kGram>(g, lds);
    if (threadIdx.x < (unsigned)kGram)
        a.gram[((size_t)(date - 1) * kGram + threadIdx.x) * kLsmGramBlocks + blockIdx.x] = t;
}

// P upper triangular with P^T P = G + shift I (Cholesky; the shift, relative to the trace,
// keeps it defined for rank-deficient or extremely ill-conditioned G), with its reciprocal
// diagonal.
__device__ __forceinline__ DateFactor shifted_cholesky(const double (&g)[kGram]) {
    DateFactor f;
    const double shift = kLsmShift * (g[ridx(0, 0)] + g[ridx(1, 1)] + g[ridx(2, 2)] + g[ridx(3, 3)]);
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
        double d = g[ridx(i, i)] + shift;
#pragma unroll
        for (int k = 0; k < i; ++k) d -= f.r[ridx(k, i)] * f.r[ridx(k, i)];
        const bool live = d > 0.0;
        const double pi = live ? sqrt(d) : 0.0;
        f.r[ridx(i, i)] = pi;
        f.inv_d[i] = live ? 1.0 / pi : 0.0;
#pragma unroll
        for (int j = i + 1; j < kLsmBasis; ++j) {
            double t2 = g[ridx(i, j)];
#pragma unroll
            for (int k = 0; k < i; ++k) t2 -= f.r[ridx(k, i)] * f.r[ridx(k, j)];
            f.r[ridx(i, j)] = t2 * f.inv_d[i];
        }
    }
    return f;
}

// Grid (M - 1, runs): one block per date adds its gb blocks' Gram sums (lane b holds block
// b's) and factors them, G + shift I = P^T P. A date whose sample held fewer than
// kLsmMinSample rows (while the sample is not every path) is summed again over every path
// (one block: slow, but only for such thin dates). One Cholesky step suffices: for rows of
// condition number k, the preconditioned rows' condition number is about 1 + k^2 1e-16 (1.0004
// at k = 2e6), and the date kernel's normal equations in them lose only its square times
// 1e-16; beyond k ~ 1e8 the shift keeps P defined.
template <bool FP64>
__global__ __launch_bounds__(kLsmLanes) void lsm_gram_merge_kernel(LsmArgs a, uint32_t gb) {
    select_run(a);
    const int date = (int)blockIdx.x + 1;
    if (date == 1 && threadIdx.x < kLsmSyncGroups) a.sync[threadIdx.x * kLsmSyncWords] = 0; // before use
    __shared__ double lds[kLsmLdsDoubles];
    double g[kGram];
#pragma unroll
    for (int k = 0; k < kGram; ++k)
        g[k] = threadIdx.x < gb ? a.gram[((size_t)(date - 1) * kGram + k) * kLsmGramBlocks + threadIdx.x]
                                : 0.0;
    double* res = lds + kLsmLdsRes;
    double t = block_sums<kGram>(g, lds);
    if (threadIdx.x == 10) res[kNe + 4] = t; // the row count, in a slot nothing else writes
    __syncthreads();
    const bool thin = res[kNe + 4] < kLsmMinSample && a.sample_n < a.n_paths; // uniform
    if (thin) {
#pragma unroll
        for (int k = 0; k < kGram; ++k) g[k] = 0.0;
        gram_rows<FP64>(a, date, 1, 0, (uint32_t)((a.n_paths + kLsmLanes - 1) / kLsmLanes), a.n_paths, g);
        t = block_sums<kGram>(g, lds);
    }
    if (threadIdx.x < (unsigned)kGram) res[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int k = 0; k < kGram; ++k) g[k] = res[k];
    a.factors[date] = shifted_cholesky(g);
}

// ---- One date of the backward induction --------------------------------------------------
__device__ __forceinline__ double uniform_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// A date's preconditioner (the same for every lane) held in scalar registers.
__device__ __forceinline__ DateFactor uniform_factor(const DateFactor& g) {
    DateFactor f;
#pragma unroll
    for (int i = 0; i < 10; ++i) f.r[i] = uniform_f64(g.r[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) f.inv_d[i] = uniform_f64(g.inv_d[i]);
    return f;
}

// A date's coefficients from its normal-equation sums (G' packed upper, q) and preconditioner
// P: the least-squares problem in the preconditioned rows W = A P^-1 is well conditioned, so
// its normal equations are solved directly, c' = G'^-1 q by Cholesky G' = U^T U, and the
// coefficients are beta = P^-1 c' by back-substitution (linalg.cpp:58-69; a zero diagonal
// gives coefficient 0). A non-positive pivot of G' marks a dropped column: its component
// is 0. Square roots and reciprocals are the hardware estimates refined by Newton steps.
__device__ __forceinline__ void ne_solve(const double (&g)[kNe], const DateFactor& f, double* beta) {
    double u[10], iu[4];
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
        double d = g[ridx(i, i)];
#pragma unroll
        for (int k = 0; k < i; ++k) d = __builtin_fma(-u[ridx(k, i)], u[ridx(k, i)], d);
        const bool live = d > 0.0;
        iu[i] = live ? fast_rsqrt_f64(d) : 0.0;
#pragma unroll
        for (int j = i + 1; j < kLsmBasis; ++j) {
            double t = g[ridx(i, j)];
#pragma unroll
            for (int k = 0; k < i; ++k) t = __builtin_fma(-u[ridx(k, i)], u[ridx(k, j)], t);
            u[ridx(i, j)] = t * iu[i];
        }
    }
    double t[4], c[4];
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) { // U^T t = q
        double s = g[10 + i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = __builtin_fma(-u[ridx(k, i)], t[k], s);
        t[i] = s * iu[i];
    }
#pragma unroll
    for (int i = kLsmBasis - 1; i >= 0; --i) { // U c' = t
        double s = t[i];
#pragma unroll
        for (int j = i + 1; j < kLsmBasis; ++j) s = __builtin_fma(-u[ridx(i, j)], c[j], s);
        c[i] = s * iu[i];
    }
#pragma unroll
    for (int i = kLsmBasis - 1; i >= 0; --i) { // P beta = c'
        double s = c[i];
#pragma unroll
        for (int j = i + 1; j < kLsmBasis; ++j) s = __builtin_fma(-f.r[ridx(i, j)], beta[j], s);
        beta[i] = s * f.inv_d[i];
    }
}

// One date of the backward induction. With fit, the later date's regression is solved here:
// wave 0 of every block adds the later date's block sums (written by the previous launch, in
// the other half of the double-buffered sums) in the same fixed order, so every block holds
// the same coefficients, and block 0 records them; no separate solve launch per date. Without
// fit, the policy in coeffs/has is applied. CAP: the regression captures the network's
// training samples (a.cap_*).
template <bool FP64, bool CAP>
__global__ __launch_bounds__(kLsmLanes) void lsm_date_kernel(LsmArgs a, int date, int regress,
                                                             int fit) {
    select_run(a);
    __shared__ double lds[kLsmLdsDoubles];
    const int later = date + 1;
    const bool solve = fit && later <= a.m - 1; // uniform
    // Every load is issued up front, those the solve needs first: wave 0's block sums (lane l:
    // blocks l, l + 64, l + 128, l + 192; the buffer is zero past the blocks), the
    // preconditioner and discount powers, then the lane's paths (path 0 past the end, store
    // column 1 for the final date 0), so no load waits inside a branch and unused values are
    // ignored.
    double gv[kNe][4];
    if (solve && threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < kNe; ++k) {
            const double* src = a.ne + ((size_t)(later & 1) * kNe + k) * kLsmNeStride;
#pragma unroll
            for (int q = 0; q < 4; ++q) gv[k][q] = src[threadIdx.x + 64 * q];
        }
    }
    const DateFactor f = uniform_factor(a.factors[regress ? date : 1]);
    const double d_later = a.dpow[later <= a.m ? later : a.m];
    const double d_date = a.dpow[date];
    const int col = date >= 1 ? date : 1;
    double v[kLsmChunk], s1[kLsmChunk], s0[kLsmChunk];
    auto load_chunk = [&](uint32_t j0) {
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
            const uint64_t q = (j0 + u < a.per_lane && p < a.n_paths) ? p : 0;
            v[u] = a.val[q];
            s1[u] = store_get<FP64>(a, later, q);
            s0[u] = store_get<FP64>(a, col, q);
        }
    };
    load_chunk(0);
    double beta[4] = {0.0, 0.0, 0.0, 0.0};
    bool use = false;
    if (later <= a.m - 1) { // uniform
        if (solve) {
            // Wave 0 adds the blocks' sums as (b0 + b1) + (b2 + b3) per lane, then over its
            // lanes, solves, and hands the coefficients to the other waves through LDS; one
            // block barrier.
            double* res = lds + kLsmLdsRes; // totals [kNe], then beta [4] and the flag
            if (threadIdx.x < 64) {
                double g[kNe];
#pragma unroll
                for (int k = 0; k < kNe; ++k) g[k] = (gv[k][0] + gv[k][1]) + (gv[k][2] + gv[k][3]);
                const double t = wave_sums<kNe>(g, lds);
                if ((threadIdx.x & 3) == 0 && (threadIdx.x >> 2) < (unsigned)kNe) res[threadIdx.x >> 2] = t;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                double tot[kNe];
#pragma unroll
                for (int k = 0; k < kNe; ++k) tot[k] = res[k];
                const bool u = tot[14] >= (double)kLsmBasis; // lsm.cpp:102-104: >= 4 ITM rows
                if (u) {
                    // The sums regress the values discounted to t = 0; the later date's target
                    // is those over disc^later.
                    ne_solve(tot, a.factors[later], beta);
                    const double inv = 1.0 / d_later;
#pragma unroll
                    for (int j = 0; j < 4; ++j) beta[j] *= inv;
                }
                if (threadIdx.x == 0) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) res[kNe + j] = beta[j];
                    res[kNe + 4] = u ? 1.0 : 0.0;
                    if (blockIdx.x == 0) {
                        a.has[later] = u ? 1 : 0;
                        if (u)
                            for (int j = 0; j < 4; ++j) a.coeffs[4 * later + j] = beta[j];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 4; ++j) beta[j] = res[kNe + j];
            use = res[kNe + 4] != 0.0;
        } else {
            use = a.has[later] != 0;
            if (use)
                for (int j = 0; j < 4; ++j) beta[j] = a.coeffs[4 * later + j];
        }
    }
    const ContPoly cp = cont_poly(beta);
    double acc[kNe];
#pragma unroll
    for (int k = 0; k < kNe; ++k) acc[k] = 0.0;
    // Rows are summed in ascending path order, as one path at a time would.
    for (uint32_t j0 = 0; j0 < a.per_lane; j0 += kLsmChunk) {
        if (j0 > 0) load_chunk(j0);
        if (use) { // uniform: the decisions at the later date, screened without branches
            bool ex[kLsmChunk], unsure = false;
#pragma unroll
            for (uint32_t u = 0; u < kLsmChunk; ++u) unsure |= decide_screen(a, s1[u], cp, &ex[u]);
            if (__builtin_amdgcn_ballot_w64(unsure) != 0) { // rare: some path near the boundary
#pragma unroll
                for (uint32_t u = 0; u < kLsmChunk; ++u) {
                    bool o;
                    if (decide_screen(a, s1[u], cp, &o)) ex[u] = decide(a, s1[u], beta);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kLsmChunk; ++u) {
                const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
                if (ex[u] && j0 + u < a.per_lane && p < a.n_paths) { // exercise at the later date
                    v[u] = intrinsic_v(s1[u], a.strike, a.type) * d_later;
                    a.val[p] = v[u];
                }
            }
        }
        if (regress) { // uniform across the grid
            // Out-of-the-money and past-the-end paths are zero rows and add nothing.
#pragma unroll
            for (uint32_t u = 0; u < kLsmChunk; ++u) {
                const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
                const bool in = j0 + u < a.per_lane && p < a.n_paths;
                double row[4], w[4];
                const bool itm = design_row(a, in, s0[u], row);
                if constexpr (CAP) {
                    if (in) capture_row(a, date, p, itm, s0[u], v[u] / d_date);
                }
                const double y = itm ? v[u] : 0.0; // discounted to t = 0
                precondition_row(f, row, w);
#pragma unroll
                for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
                    for (int j = i; j < kLsmBasis; ++j)
                        acc[ridx(i, j)] = __builtin_fma(w[i], w[j], acc[ridx(i, j)]);
                    acc[10 + i] = __builtin_fma(w[i], y, acc[10 + i]);
                }
                acc[14] += itm ? 1.0 : 0.0;
            }
        }
    }
    if (!regress) return; // uniform across the grid
    const double t = block_sums<kNe>(acc, lds);
    if (threadIdx.x < (unsigned)kNe)
        a.ne[((size_t)(date & 1) * kNe + threadIdx.x) * kLsmMaxBlocks + blockIdx.x] = t;
}

// ---- Every date in one launch -------------------------------------------------------------
// The dates' kernels' work in one launch whose blocks are all resident at once (the host
// checks the occupancy before choosing it, and chains such launches so that only one occupies
// the device at a time). A block is one control wave and kLsmWaves path waves. Path lane t (of
// kLsmLanes) holds the paths the dates' kernels give lane t of the block (at most kLsmChunk:
// n <= 2^20 per run) and keeps their values and the weights e^{-x/2} of the date's rows in
// registers; the store columns come into LDS by DMA, each once, a date ahead. The blocks
// exchange each date's normal-equation sums through the ne buffer with device-scope (sc1)
// accesses and no cache write-back or invalidation: the control wave stores the block's 15
// sums at device scope, waits until the stores are complete, then adds one to its group's
// counter (block b, counter b mod kLsmSyncGroups, so that the additions of one date contend in
// small groups); a block needs the later date's sums once every counter holds its blocks'
// additions for that date, and its control wave then reads all blocks' sums into LDS by
// device-scope 16-byte DMA loads, adds them, solves and hands the coefficients to the path
// waves (one barrier); the path waves' wave sums of the date come back through LDS (a second
// barrier). Every sum is formed by the same operations in the same order as in the dates'
// kernels, so the coefficients, decisions and values are bit for bit theirs. The path waves
// form the next date's rows (weights, Gram sums and count) while the control wave waits, and
// rebuild each row from its weight for the q sums. A wait gives up after kLsmSpinLimit polls
// (seconds; never reached while every block is resident): the block stops waiting and its
// values become NaN, so the run's price reads NaN and the entry point reports a device error.
constexpr int kLsmResidentMaxDates = 512; // discount powers held in LDS
constexpr uint32_t kLsmSpinLimit = 1u << 21;
constexpr int kLsmResidentLanes = kLsmLanes + 64; // the control wave and the path waves
constexpr int kLsmStageDoubles = kNe * kLsmNeStride;
constexpr int kGramSums = 11; // G' packed (sums 0..9) and the count (sum 14)
constexpr int kLsmPathRegion = kGramSums * kLsmWaveStride; // a path wave's wave_sums region

template <bool FP64>
using StoreT = typename std::conditional<FP64, double, float>::type;

// Cache policy of a device-scope (sc1) access: coherent across the device's XCDs without a
// write-back or invalidation of the L2.
constexpr int kCpolSc1 = 16;
// s_waitcnt with at most n vector-memory operations outstanding (expcnt, lgkmcnt not waited).
constexpr int vmcnt_at_most(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// A path wave's copy of one store column in LDS, filled by 4-byte DMA loads so no register is
// held while they are in flight: cb[u][lane] = path u of the lane (FP32), or its low and high
// words at cb[2u][lane] and cb[2u + 1][lane] (FP64); path 0 past the end. The lane's path u is
// at byte boff + u stride of the column.
template <bool FP64>
constexpr int kColWords = (FP64 ? 2 : 1) * (int)kLsmChunk * 64;
template <bool FP64>
constexpr int kColLoads = (FP64 ? 2 : 1) * (int)kLsmChunk; // DMA loads per column and wave

template <bool FP64>
__device__ __forceinline__ void stage_column(const LsmArgs& a, int date, uint32_t boff,
                                             uint32_t stride, uint32_t in_mask, uint32_t* cb) {
    const char* col = (const char*)a.store + (uint64_t)(date - 1) * a.n_paths * sizeof(StoreT<FP64>);
#pragma unroll
    for (uint32_t u = 0; u < kLsmChunk; ++u) {
        const char* src = col + (((in_mask >> u) & 1u) ? boff + u * stride : 0u);
#pragma unroll
        for (uint32_t h = 0; h < (FP64 ? 2u : 1u); ++h)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + 4 * h),
                                             (__attribute__((address_space(3))) void*)(cb + ((FP64 ? 2 : 1) * u + h) * 64),
                                             4, 0, 0);
    }
}

template <bool FP64>
__device__ __forceinline__ double staged(const uint32_t* cb, uint32_t u, uint32_t l) {
    if constexpr (FP64) {
        const uint64_t lo = cb[(2 * u) * 64 + l], hi = cb[(2 * u + 1) * 64 + l];
        return __builtin_bit_cast(double, (hi << 32) | lo);
    } else {
        return (double)__builtin_bit_cast(float, cb[u * 64 + l]);
    }
}

// A date's preconditioned row of path u of the lane from its weight e^{-x/2} (0 out of the
// money and past the end) and its staged column: the dates' kernel's design_row and
// precondition_row, operation for operation.
template <bool FP64>
__device__ __forceinline__ void resident_row(const LsmArgs& a, const DateFactor& f, const uint32_t* cb,
                                             uint32_t u, double e, double* w) {
    double row[4];
    laguerre4_weighted(staged<FP64>(cb, u, threadIdx.x & 63) * a.inv_strike, e, row);
    precondition_row(f, row, w);
}

// A date's rows from its staged column: the weights e^{-x/2} of the lane's paths (0 out of the
// money and past the end; the rows follow from them and the column) and the wave's sums of what
// the rows fix alone, G' and the count, as sums 0..9 and 14 of the wave's totals `tot` (the q
// sums follow once the values are known). Per sum, the operations and their order are those of
// the dates' kernel.
template <bool FP64>
__device__ __forceinline__ void resident_rows(const LsmArgs& a, const DateFactor& f, const uint32_t* cb,
                                              uint32_t in_mask, double (&ew)[kLsmChunk],
                                              uint32_t& itm_mask, double* wl, double* tot) {
    const uint32_t l = threadIdx.x & 63;
    double g[kGramSums];
#pragma unroll
    for (int k = 0; k < kGramSums; ++k) g[k] = 0.0;
    itm_mask = 0;
#pragma unroll
    for (uint32_t u = 0; u < kLsmChunk; ++u) {
        const double s0 = staged<FP64>(cb, u, l);
        const bool itm = ((in_mask >> u) & 1u) && intrinsic_v(s0, a.strike, a.type) > 0.0;
        ew[u] = itm ? exp(-(s0 * a.inv_strike) / 2.0) : 0.0;
        double w[4];
        resident_row<FP64>(a, f, cb, u, ew[u], w);
#pragma unroll
        for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
            for (int j = i; j < kLsmBasis; ++j) g[ridx(i, j)] = __builtin_fma(w[i], w[j], g[ridx(i, j)]);
        }
        g[10] += itm ? 1.0 : 0.0;
        itm_mask |= (itm ? 1u : 0u) << u;
    }
    const double s = wave_sums<kGramSums>(g, wl);
    if ((l & 3) == 0 && (l >> 2) < (unsigned)kGramSums) tot[(l >> 2) < 10 ? (l >> 2) : 14] = s;
}

// A date's preconditioner staged in LDS (slot of 16 doubles) by a 4-byte DMA load per lane of
// lanes 0..27.
__device__ __forceinline__ void stage_factor(const DateFactor* src, double* slot) {
    static_assert(sizeof(DateFactor) == 28 * 4, "28 words");
    if ((threadIdx.x & 63) < 28)
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)((const uint32_t*)src + (threadIdx.x & 63)),
            (__attribute__((address_space(3))) void*)slot, 4, 0, 0);
}
__device__ __forceinline__ DateFactor staged_factor(const double* slot) {
    DateFactor f;
#pragma unroll
    for (int i = 0; i < 10; ++i) f.r[i] = uniform_f64(slot[i]);
#pragma unroll
    for (int i = 0; i < 4; ++i) f.inv_d[i] = uniform_f64(slot[10 + i]);
    return f;
}

template <bool FP64, bool CAP>
__global__ __launch_bounds__(kLsmResidentLanes) void lsm_resident_kernel(LsmArgs a, int fit) {
    select_run(a);
    __shared__ double regions[kLsmWaveRegion + kLsmWaves * kLsmPathRegion]; // wave_sums regions
    __shared__ uint32_t cols[kLsmWaves][2][kColWords<FP64>]; // column d at [d & 1], per path wave
    __shared__ double tot[2][kLsmWaves][kNe]; // the path waves' totals of the even / odd dates
    __shared__ double res[kNe + 5];           // the solve's totals, then beta [4] and the flag
    __shared__ double stage[kLsmStageDoubles]; // [sum][block]: the later date's block sums
    __shared__ double dpw[kLsmResidentMaxDates + 1];
    __shared__ double fsl[4][16];             // date d's preconditioner at [d & 3]
    __shared__ int dead_flag;
    const int m = a.m;
    const uint32_t blk = blockIdx.x;
    if (threadIdx.x < 64) { // the control wave
        const uint32_t l = threadIdx.x;
        if (l == 0) {
            double d = 1.0; // disc^k by repeated multiplication, as the dpow table
            for (int k = 0; k <= m; ++k) {
                dpw[k] = d;
                d *= a.disc;
            }
            dead_flag = 0;
        }
        if (fit && blk == 0) { // this run's policy, cleared before the first date records one
            for (int i = (int)l; i <= m; i += 64) {
                a.has[i] = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) a.coeffs[4 * i + j] = 0.0;
            }
        }
        // blocks adding to this lane's counter
        const uint32_t grp_blocks = l < a.blocks ? (a.blocks - l + kLsmSyncGroups - 1) / kLsmSyncGroups : 0;
        if (fit && m >= 2) stage_factor(a.factors + (m - 1), fsl[(m - 1) & 3]);
        __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0));
        __syncthreads(); // the discount powers, the flag, the cleared policy, a factor
        bool dead = false; // a wait of this block gave up
        for (int date = m - 1; date >= 0; --date) {
            const int later = date + 1;
            // Date - 1's preconditioner, for the path waves' rows after this date's sums.
            if (fit && date >= 2) stage_factor(a.factors + (date - 1), fsl[(date - 1) & 3]);
            if (fit && later <= m - 1) {
                // Every block's sums of the later date (each block adds one per date, dates
                // m - 1 down), staged in LDS: row k's 256 sums are two 16-byte loads per lane.
                // Lane g < kLsmSyncGroups reads counter g, whose blocks b = g mod kLsmSyncGroups
                // have each added once per date.
                const uint32_t target = (uint32_t)(m - later) * grp_blocks;
                uint32_t spins = 0;
                while (!dead) {
                    const uint32_t c = l < kLsmSyncGroups
                                           ? __hip_atomic_load(a.sync + l * kLsmSyncWords, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT)
                                           : target;
                    if (__builtin_amdgcn_ballot_w64(c < target) == 0) break;
                    if (++spins > kLsmSpinLimit) dead = true;
                    __builtin_amdgcn_s_sleep(1);
                }
                // Device-scope (sc1) reads of the sums, issued once the count is seen.
                const double* src = a.ne + (size_t)(later & 1) * kNe * kLsmNeStride;
#pragma unroll
                for (int k = 0; k < kNe; ++k) {
#pragma unroll
                    for (int h = 0; h < kLsmNeStride / 128; ++h)
                        __builtin_amdgcn_global_load_lds(
                            (const __attribute__((address_space(1))) void*)(src + k * kLsmNeStride + h * 128 + 2 * l),
                            (__attribute__((address_space(3))) void*)(stage + k * kLsmNeStride + h * 128),
                            16, 0, kCpolSc1);
                }
                __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0)); // the staged sums (and factor)
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // As the dates' kernel: lane l adds blocks (l + l64) + (l128 + l192), zero past
                // the blocks, then the wave adds its lanes.
                double g[kNe];
#pragma unroll
                for (int k = 0; k < kNe; ++k) {
                    double q4[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        q4[q] = l + 64 * q < a.blocks ? stage[k * kLsmNeStride + l + 64 * q] : 0.0;
                    g[k] = (q4[0] + q4[1]) + (q4[2] + q4[3]);
                }
                const double tw = wave_sums<kNe>(g, regions);
                if ((l & 3) == 0 && (l >> 2) < (unsigned)kNe) res[l >> 2] = tw;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                double tt[kNe];
#pragma unroll
                for (int k = 0; k < kNe; ++k) tt[k] = res[k];
                const bool u = tt[14] >= (double)kLsmBasis; // lsm.cpp:102-104
                double bz[4] = {0.0, 0.0, 0.0, 0.0};
                if (u) {
                    ne_solve(tt, staged_factor(fsl[later & 3]), bz);
                    const double inv = 1.0 / dpw[later];
#pragma unroll
                    for (int j = 0; j < 4; ++j) bz[j] *= inv;
                }
                if (l == 0) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) res[kNe + j] = bz[j];
                    res[kNe + 4] = u ? 1.0 : 0.0;
                    if (dead) dead_flag = 1;
                    if (blk == 0) {
                        a.has[later] = u ? 1 : 0;
                        if (u)
                            for (int j = 0; j < 4; ++j) a.coeffs[4 * later + j] = bz[j];
                    }
                }
                __syncthreads(); // the coefficients to the path waves
            }
            if (fit && date >= 1) {
                __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0)); // date - 1's factor is in
                __syncthreads(); // the path waves' totals of this date
                // As block_sums: lane i < kNe adds sum i of the kLsmWaves waves pairwise.
                double tk[kLsmWaves];
#pragma unroll
                for (int k = 0; k < kLsmWaves; ++k) tk[k] = l < (unsigned)kNe ? tot[date & 1][k][l] : 0.0;
#pragma unroll
                for (int h = kLsmWaves / 2; h >= 1; h /= 2) {
#pragma unroll
                    for (int k = 0; k < h; ++k) tk[k] = tk[2 * k] + tk[2 * k + 1];
                }
                // Device-scope (sc1) stores of the sums, complete (the wave's store count drained)
                // before lane 0's one addition to the count.
                if (l < (unsigned)kNe)
                    __hip_atomic_store(a.ne + ((size_t)(date & 1) * kNe + l) * kLsmNeStride + blk, tk[0],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0));
                __asm__ __volatile__("" ::: "memory");
                if (l == 0)
                    __hip_atomic_fetch_add(a.sync + (blk % kLsmSyncGroups) * kLsmSyncWords, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads(); // the flag to the path waves
    } else { // a path wave
        const uint32_t t = threadIdx.x - 64, pw = t >> 6, l = t & 63; // path lane, path wave
        double* wl = regions + kLsmWaveRegion + pw * kLsmPathRegion;
        // Byte offsets of the lane's paths within a column (n <= 2^20 here): path u at
        // boff + u stride.
        const uint32_t esz = (uint32_t)sizeof(StoreT<FP64>);
        const uint32_t boff = (uint32_t)lane_path(0, blk, a.blocks, t) * esz;
        const uint32_t stride = a.blocks * (uint32_t)kLsmLanes * esz;
        uint32_t in_mask = 0;
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = lane_path(u, blk, a.blocks, t);
            in_mask |= (u < a.per_lane && p < a.n_paths ? 1u : 0u) << u;
        }
        // Column d is staged at cols[pw][d & 1] during date d + 1 (after its decisions, which
        // read the buffer's previous column d + 2), read by the rows of date d and the
        // decisions at d.
        stage_column<FP64>(a, m, boff, stride, in_mask, cols[pw][m & 1]);
        if (m >= 2) stage_column<FP64>(a, m - 1, boff, stride, in_mask, cols[pw][(m - 1) & 1]);
        __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0));
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double v[kLsmChunk];
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u)
            v[u] = intrinsic_v(staged<FP64>(cols[pw][m & 1], u, l), a.strike, a.type) * a.disc_m;
        __syncthreads(); // the discount powers, the flag, the cleared policy, a factor
        double ew[kLsmChunk]; // the date's row weights
        uint32_t itm_mask = 0;
        if (fit && m >= 2)
            resident_rows<FP64>(a, staged_factor(fsl[(m - 1) & 3]), cols[pw][(m - 1) & 1], in_mask, ew,
                                itm_mask, wl, tot[(m - 1) & 1][pw]);
        for (int date = m - 1; date >= 0; --date) {
            const int later = date + 1;
            const uint32_t* cl = cols[pw][later & 1]; // the later date's column
            double beta[4] = {0.0, 0.0, 0.0, 0.0};
            bool use = false;
            if (later <= m - 1) { // uniform
                if (fit) {
                    __syncthreads(); // the control wave's coefficients
#pragma unroll
                    for (int j = 0; j < 4; ++j) beta[j] = uniform_f64(res[kNe + j]);
                    use = res[kNe + 4] != 0.0;
                } else {
                    use = a.has[later] != 0;
                    if (use)
                        for (int j = 0; j < 4; ++j) beta[j] = a.coeffs[4 * later + j];
                }
            }
            if (use) { // uniform: the decisions at the later date, as the dates' kernel
                const double d_later = dpw[later];
                const ContPoly cp = cont_poly(beta);
                uint32_t ex = 0, unsure = 0;
#pragma unroll
                for (uint32_t u = 0; u < kLsmChunk; ++u) {
                    bool e;
                    unsure |= (decide_screen(a, staged<FP64>(cl, u, l), cp, &e) ? 1u : 0u) << u;
                    ex |= (e ? 1u : 0u) << u;
                }
                if (__builtin_amdgcn_ballot_w64(unsure != 0) != 0) { // rare: a path near the boundary
#pragma unroll 1
                    for (uint32_t u = 0; u < kLsmChunk; ++u)
                        if ((unsure >> u) & 1u)
                            ex = (ex & ~(1u << u)) | ((decide(a, staged<FP64>(cl, u, l), beta) ? 1u : 0u) << u);
                }
#pragma unroll
                for (uint32_t u = 0; u < kLsmChunk; ++u)
                    if ((ex >> u) & 1u) v[u] = intrinsic_v(staged<FP64>(cl, u, l), a.strike, a.type) * d_later;
            }
            // The decisions have read the later column: its buffer takes the column two dates
            // earlier (the rows of date - 1, the decisions of date - 2).
            if (date >= 2) stage_column<FP64>(a, date - 1, boff, stride, in_mask, cols[pw][(date - 1) & 1]);
            if (fit && date >= 1) { // uniform: this date's q sums
                const DateFactor fd = staged_factor(fsl[date & 3]);
                double q[kLsmBasis] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (uint32_t u = 0; u < kLsmChunk; ++u) {
                    double w[4];
                    resident_row<FP64>(a, fd, cols[pw][date & 1], u, ew[u], w);
                    const bool itm = (itm_mask >> u) & 1u;
                    if constexpr (CAP) {
                        if ((in_mask >> u) & 1u) {
                            const uint64_t p = lane_path(u, blk, a.blocks, t);
                            capture_row(a, date, p, itm, store_get<FP64>(a, date, p), v[u] / dpw[date]);
                        }
                    }
                    const double y = itm ? v[u] : 0.0; // discounted to t = 0
#pragma unroll
                    for (int i = 0; i < kLsmBasis; ++i) q[i] = __builtin_fma(w[i], y, q[i]);
                }
                const double sq = wave_sums<kLsmBasis>(q, wl);
                if ((l & 3) == 0 && (l >> 2) < (unsigned)kLsmBasis) tot[date & 1][pw][10 + (l >> 2)] = sq;
                __syncthreads(); // this date's totals to the control wave
            }
            if (fit && date >= 2) { // the next date's rows, while the control wave waits
                __builtin_amdgcn_s_waitcnt(vmcnt_at_most(0)); // its column
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                resident_rows<FP64>(a, staged_factor(fsl[(date - 1) & 3]), cols[pw][(date - 1) & 1], in_mask,
                                    ew, itm_mask, wl, tot[(date - 1) & 1][pw]);
            }
        }
        __syncthreads(); // the control wave's flag
        const bool lost = dead_flag != 0;
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u)
            if ((in_mask >> u) & 1u) a.val[lane_path(u, blk, a.blocks, t)] = lost ? __builtin_nan("") : v[u];
    }
}

// Welford over the final cash flows in the canonical tile order.
template <bool FP64>
__global__ __launch_bounds__(kTileLanes) void lsm_welford_kernel(LsmArgs a) {
    select_run(a);
    __shared__ Welford lds4[4];
    const uint64_t tile = blockIdx.x;
    Welford w = {0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPathsPerLane; ++k) {
        const uint64_t p = tile * kTilePaths + (uint64_t)(k * kTileLanes) + threadIdx.x;
        if (p < a.n_paths) {
            const double x = a.val[p]; // discounted to t = 0 (lsm.cpp:125-127)
            w.n += 1;
            const double d = x - w.mean;
            w.mean += d * (1.0 / (double)(k + 1));
            w.m2 += d * (x - w.mean);
        }
    }
    const Welford t = block_merge_256(w, lds4);
    if (threadIdx.x == 0) a.partials[tile] = t;
}

// ---- Streaming exercise by a policy (no path store) --------------------------------------
// The network's continuation value enters as a table over S/K per exercise date: it is the
// network evaluated (by the MFMA forward kernel) on a uniform grid, read with linear
// interpolation. Outside the table's range the QR coefficients decide (the fallback).
// A policy run's merged results, adjacent so that one copy brings them back: the network's
// and the QR policy's Welford totals and the counts of joint ITM decisions, disagreements and
// table-range fallbacks.
struct PolicyResults {
    Welford nn, qr;
    unsigned long long ctr[3];
};

struct PolicyArgs {
    const float* table;           // [date 1..M-1][grid], kTargetScale * continuation / K
    uint32_t grid;
    double xlo, xhi, inv_h;       // S/K range of the table and 1/spacing
    double inv_strike;            // 1/K for the table lookup
    double cont_scale;            // K / kTargetScale: table units to cash
    Welford* part_nn;             // [tiles + 1]
    Welford* part_qr;             // [tiles + 1]
    uint32_t* tile_counts;        // [3][tiles] ITM decisions, disagreements, fallbacks
    int judge;                    // 0: the QR policy only decides outside the table range;
                                  // 1: both policies on every path; 2: the QR policy alone
    // FP32 decisions (lsm_serve_kernel): table coordinate u = s ua + ub, in range for
    // 0 <= u <= umax = grid - 1, cell index clamped to gmax = grid - 2; in the money iff
    // phi s > kedge (phi = +1 call, -1 put; kedge = phi times the strike rounded outward to
    // FP32); exercise margin (phi s + phik) - c cont_scale_f (phik = -phi K).
    float ua, ub, umax, gmax;
    float kedge, phik, cont_scale_f;
};

// Both policies' state along one path: stopped flags, exercise dates and values, and the
// path's counts of joint ITM decisions, disagreements and table-range fallbacks.
struct PolicyState {
    bool n_done, q_done;
    int n_date, q_date;
    double n_val, q_val;
    uint32_t dec, dis, fb;
};

// has: the coefficient flags [date] in LDS (a byte load from global memory would be a
// vector load and a wait on every visit).
__device__ __forceinline__ void policy_visit(const LsmArgs& a, const PolicyArgs& q,
                                             const uint8_t* has, PolicyState& ps, int date,
                                             double s) {
    const double ex = intrinsic_v(s, a.strike, a.type);
    if (date == a.m) { // lsm.cpp:76-81
        if (!ps.n_done) { ps.n_done = true; ps.n_date = date; ps.n_val = ex; }
        if (!ps.q_done) { ps.q_done = true; ps.q_date = date; ps.q_val = ex; }
        return;
    }
    if (!(ex > 0.0)) return;
    const double x = s * q.inv_strike;
    const bool in_range = x >= q.xlo && x <= q.xhi;
    const bool both = !ps.n_done && !ps.q_done;
    bool dq = false;
    if ((!ps.q_done || both || !in_range) && has[date]) { // lsm.cpp:110-122
        double b0, b1, b2, b3;
        {
            double b[4];
            laguerre4(s * a.inv_strike, b);
            b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
        }
        const double* beta = a.coeffs + 4 * date;
        double cont = 0.0;
        cont += beta[0] * b0;
        cont += beta[1] * b1;
        cont += beta[2] * b2;
        cont += beta[3] * b3;
        dq = ex > cont;
    }
    bool dn = false;
    if (!ps.n_done) {
        if (in_range) {
            const float u = (float)((x - q.xlo) * q.inv_h);
            uint32_t g = (uint32_t)u;
            if (g > q.grid - 2) g = q.grid - 2;
            const float* row = q.table + (uint64_t)(date - 1) * q.grid;
            const float lo = row[g], hi = row[g + 1];
            const float c = __builtin_fmaf(u - (float)g, hi - lo, lo);
            dn = ex > (double)c * q.cont_scale;
        } else {
            dn = dq;
            ps.fb += 1;
        }
    }
    if (both) {
        ps.dec += 1;
        ps.dis += (dn != dq) ? 1u : 0u;
    }
    if (!ps.q_done && dq) { ps.q_done = true; ps.q_date = date; ps.q_val = ex; }
    if (!ps.n_done && dn) { ps.n_done = true; ps.n_date = date; ps.n_val = ex; }
}

// The tile's close for a policy kernel (one path per thread, a block of 1024 threads is one
// canonical tile: thread t computed path tile * 1024 + t): the per-tile counts of ITM
// decisions, disagreements and fallbacks, and the Welford partials of the policies formed (NN:
// the network's, QR: the QR policy's) rebuilt from LDS in the canonical order (lane l < 256
// adds paths l, l+256, l+512, l+768 in turn, then the 256-lane tree in waves 0..3), so they
// equal the partials of the 256-lane path kernels.
template <bool NN, bool QR>
__device__ __forceinline__ void policy_tile_close(const LsmArgs& a, const PolicyArgs& q,
                                                  uint64_t tile, uint32_t t, double cn, double cq,
                                                  uint32_t c_dec, uint32_t c_dis, uint32_t c_fb) {
    __shared__ double sh_n[NN ? kTilePaths : 1], sh_q[QR ? kTilePaths : 1];
    __shared__ Welford lds_n[4], lds_q[4];
    __shared__ uint32_t sh_cnt[3][16];
    if constexpr (NN) sh_n[t] = cn;
    if constexpr (QR) sh_q[t] = cq;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        c_dec += __shfl_xor(c_dec, off, 64);
        c_dis += __shfl_xor(c_dis, off, 64);
        c_fb += __shfl_xor(c_fb, off, 64);
    }
    if ((t & 63) == 0) {
        sh_cnt[0][t >> 6] = c_dec;
        sh_cnt[1][t >> 6] = c_dis;
        sh_cnt[2][t >> 6] = c_fb;
    }
    __syncthreads();
    if (t < 3) { // per-tile counts, summed by lsm_policy_close_kernel (no same-address atomics)
        uint32_t sum = 0;
        for (int w = 0; w < 16; ++w) sum += sh_cnt[t][w];
        q.tile_counts[(uint64_t)t * gridDim.x + tile] = sum;
    }
    if (t < (uint32_t)kTileLanes) { // waves 0..3 (wave-uniform)
        Welford wn = {0, 0.0, 0.0}, wq = {0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < kPathsPerLane; ++k) {
            const uint32_t i = (uint32_t)(k * kTileLanes) + t;
            if (tile * kTilePaths + i >= a.n_paths) break;
            const double rk = 1.0 / (double)(k + 1);
            if constexpr (NN) {
                const double xn = sh_n[i];
                wn.n += 1;
                const double d = xn - wn.mean;
                wn.mean += d * rk;
                wn.m2 += d * (xn - wn.mean);
            }
            if constexpr (QR) {
                const double xq = sh_q[i];
                wq.n += 1;
                const double d = xq - wq.mean;
                wq.mean += d * rk;
                wq.m2 += d * (xq - wq.mean);
            }
        }
        // the 256-lane tree of block_merge_256
        const int lane = t & 63;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            if constexpr (NN) {
                Welford on;
                on.n = __shfl_down(wn.n, off, 64);
                on.mean = __shfl_down(wn.mean, off, 64);
                on.m2 = __shfl_down(wn.m2, off, 64);
                if ((lane & (2 * off - 1)) == 0) wn = welford_merge(wn, on);
            }
            if constexpr (QR) {
                Welford oq;
                oq.n = __shfl_down(wq.n, off, 64);
                oq.mean = __shfl_down(wq.mean, off, 64);
                oq.m2 = __shfl_down(wq.m2, off, 64);
                if ((lane & (2 * off - 1)) == 0) wq = welford_merge(wq, oq);
            }
        }
        if (lane == 0) {
            lds_n[t >> 6] = wn;
            lds_q[t >> 6] = wq;
        }
    }
    __syncthreads();
    if (t == 0) {
        if constexpr (NN)
            q.part_nn[tile] = welford_merge(welford_merge(lds_n[0], lds_n[1]),
                                            welford_merge(lds_n[2], lds_n[3]));
        if constexpr (QR)
            q.part_qr[tile] = welford_merge(welford_merge(lds_q[0], lds_q[1]),
                                            welford_merge(lds_q[2], lds_q[3]));
    }
}

// A policy run's close in one launch: the fixed-order merges of the policies formed (mode 0:
// the network's; 1: both; 2: the QR policy's), one block each, and a last block summing the
// per-tile counts into res->ctr.
__global__ __launch_bounds__(kMergeGroups) void lsm_policy_close_kernel(
    const Welford* __restrict__ pn, const Welford* __restrict__ pq, uint64_t tiles,
    const uint32_t* __restrict__ tile_counts, int mode, PolicyResults* __restrict__ res) {
    __shared__ Welford lds[kMergeGroups];
    __shared__ unsigned long long part[3][kMergeGroups / 64];
    if (blockIdx.x + 1 == gridDim.x) {
        unsigned long long acc[3] = {0, 0, 0};
        for (uint64_t i = threadIdx.x; i < tiles; i += blockDim.x)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] += tile_counts[(uint64_t)c * tiles + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
            if ((threadIdx.x & 63) == 0) part[c][threadIdx.x >> 6] = acc[c];
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            unsigned long long sum = 0;
            for (int w = 0; w < kMergeGroups / 64; ++w) sum += part[threadIdx.x][w];
            res->ctr[threadIdx.x] = sum;
        }
        return;
    }
    const bool qr = mode == 2 || blockIdx.x == 1;
    const Welford total = merge_tiles_fixed_order(qr ? pq : pn, tiles, lds);
    if (threadIdx.x == 0) (qr ? res->qr : res->nn) = total;
}

// FP64 mode: each path regenerated exactly as the store kernel writes it (the reference's
// arithmetic) and stopped at each policy's first exercise date; the cash flow is then
// discounted by one multiplication per date, the order backward induction applies them in,
// so the QR policy reproduces the backward-induction price bit for bit.
__global__ __launch_bounds__(4 * kTileLanes) void lsm_policy_f64_kernel(LsmArgs a, PolicyArgs q) {
    extern __shared__ uint8_t sh_has[]; // [m + 1], launched with that much dynamic LDS
    const uint64_t tile = blockIdx.x;
    const uint32_t t = threadIdx.x;
    for (int i = (int)t; i <= a.m; i += (int)blockDim.x) sh_has[i] = a.has[i];
    __syncthreads();
    const uint64_t p = tile * kTilePaths + t;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    // Without the judge the QR policy starts stopped: it then only decides for the network
    // outside the table range, and its own price is not formed. Serving the QR policy alone,
    // the network starts stopped and the table is never read.
    PolicyState ps = {q.judge == 2, q.judge == 0, 0, 0, 0.0, 0.0, 0u, 0u, 0u};
    if (p < a.n_paths) {
        // t < 1024 = kTilePaths: the block's paths share p >> 32, a scalar here.
        const uint32_t plo = (uint32_t)p, phh = (uint32_t)((tile * kTilePaths) >> 32);
        double sp = a.spot;
        for (int step = 0; step < a.m && !(ps.n_done && ps.q_done); ++step) {
            const double z = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
            sp = sp * exp(a.drift + a.vsd * z);
            policy_visit(a, q, sh_has, ps, step + 1, sp);
        }
    }
    const double cn = ps.n_val * disc_pow(a.disc, ps.n_date), cq = ps.q_val * disc_pow(a.disc, ps.q_date);
    policy_tile_close<true, true>(a, q, tile, t, cn, cq, ps.dec, ps.dis, ps.fb);
}

// FP32 mode, one instance per judge mode (0: the network, the QR policy deciding only outside
// the table range; 1: both policies; 2: the QR policy alone) and option type (PUT: type 1).
// The path is the store kernel's
// FP32 log2-space path. The network's decision is taken in FP32 without branches: the
// in-the-money test against the strike rounded outward to FP32 (exactly the FP64 test
// intrinsic > 0), the table coordinate u = s ua + ub by one FMA (range [0, grid - 1]), the
// interpolated continuation c and the exercise margin phi (s - K) - c K / 10 by FMAs; every
// date's table entries are loaded (the index clamped into the table, used only in range), so
// the four dates of a draw block issue their loads before the first decision waits. The QR
// policy's decisions are policy_visit's FP64 ones: screened by the date kernel's estimate with
// the exact FP64 decision for the lanes it leaves unsure (the judge and the QR-alone mode), or
// exact behind a branch (the table-range fallback), so its price equals the FP64-decision
// replays bit for bit. A stopped
// policy keeps its date and FP32 spot; its cash flow is formed once at the end from the FP64
// intrinsic and the discount powers backward induction applies.
template <int MODE, bool PUT>
__global__ __launch_bounds__(4 * kTileLanes, 8) void lsm_serve_kernel(LsmArgs a, PolicyArgs q) {
    extern __shared__ uint8_t sh_has[]; // [m + 1], launched with that much dynamic LDS
    const uint64_t tile = blockIdx.x;
    const uint32_t t = threadIdx.x;
    for (int i = (int)t; i <= a.m; i += (int)blockDim.x) sh_has[i] = a.has[i];
    __syncthreads();
    const uint64_t p = tile * kTilePaths + t;
    bool nd = MODE == 2, qd = MODE == 0; // stopped
    int nds = 0, qds = 0;                // stop dates (0: never stopped, no cash flow)
    float ns = 0.f, qs = 0.f;            // spot at the stop
    uint32_t dec = 0, dis = 0, fb = 0;
    if (p < a.n_paths) {
        const uint32_t plo = (uint32_t)p, phh = (uint32_t)((tile * kTilePaths) >> 32);
        const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
        const float s0 = (float)a.spot, mu2 = a.mu2, sig2 = a.sig2;
        const float kedge = q.kedge, phik = q.phik, cscale = q.cont_scale_f;
        const float ua = q.ua, ub = q.ub, umax = q.umax, gmax = q.gmax;
        const uint32_t grid = q.grid;
        const char* table = (const char*)q.table;
        const int last = a.m;
        // The QR policy's FP64 decision at date d (lsm.cpp:110-122, as policy_visit).
        auto qr_exercises = [&](int date, float s) {
            const double sd = (double)s;
            const double ex = intrinsic_v(sd, a.strike, a.type);
            double b[4];
            laguerre4(sd * a.inv_strike, b);
            const double* beta = a.coeffs + 4 * date;
            double cont = 0.0;
            cont += beta[0] * b[0];
            cont += beta[1] * b[1];
            cont += beta[2] * b[2];
            cont += beta[3] * b[3];
            return ex > cont;
        };
        // The QR policy's decision at a date with coefficients, screened as the date kernel
        // screens it (decide_screen's estimate, the exact decide() for the lanes it leaves
        // unsure: the same result as qr_exercises), for the lanes that need it.
        auto qr_screened = [&](int date, float s, bool need) {
            const double* beta = a.coeffs + 4 * date;
            const ContPoly cp = cont_poly(beta);
            bool ex = false;
            const bool unsure = need && decide_screen(a, (double)s, cp, &ex);
            if (unsure) ex = decide(a, (double)s, beta);
            return need && ex;
        };
        // One exercise date d < M of the path at spot s.
        auto visit = [&](int date, float s, float lo, float hi, float u, uint32_t g) {
            // phi (s - K) > 0 in FP64: s above (call) or below (put) the outward-rounded strike
            const bool itm = PUT ? s < -kedge : s > kedge;
            const bool inr = u >= 0.f && u <= umax;
            bool dn = false;
            if constexpr (MODE != 2) {
                const float c = __builtin_fmaf(u - (float)g, hi - lo, lo);
                const float ex = PUT ? phik - s : s + phik; // phi s - phi K
                dn = __builtin_fmaf(-c, cscale, ex) > 0.f;
            }
            if constexpr (MODE == 0) {
                // the table decides in range; the QR policy outside it (the fallback)
                const bool open = itm && !nd;
                fb += (open && !inr) ? 1u : 0u;
                if (open && inr && dn) { nd = true; nds = date; ns = s; }
                if (open && !inr && sh_has[date] && qr_exercises(date, s)) {
                    nd = true;
                    nds = date;
                    ns = s;
                }
            } else if constexpr (MODE == 2) {
                if (sh_has[date] && qr_screened(date, s, itm && !qd)) { // the flag: uniform
                    qd = true;
                    qds = date;
                    qs = s;
                }
            } else {
                const bool need_q = itm && (!qd || (!nd && !inr));
                bool dq = false;
                if (sh_has[date]) dq = qr_screened(date, s, need_q); // the flag: uniform
                const bool open = itm && !nd;
                fb += (open && !inr) ? 1u : 0u;
                dn = inr ? dn : dq;
                const bool both = open && !qd;
                dec += both ? 1u : 0u;
                dis += (both && dn != dq) ? 1u : 0u;
                if (open && dn) { nd = true; nds = date; ns = s; }
                if (itm && !qd && dq) { qd = true; qds = date; qs = s; }
            }
        };
        // The table coordinate of spot s at date d and its cell's two entries (the index
        // clamped into the table so that every lane's loads are in bounds; a 32-bit byte
        // offset from the table's base).
        auto locate = [&](int date, float s, float& lo, float& hi, float& u, uint32_t& g) {
            u = __builtin_fmaf(s, ua, ub);
            g = (uint32_t)__builtin_fminf(__builtin_fmaxf(u, 0.f), gmax);
            if constexpr (MODE != 2) {
                const uint32_t off = ((uint32_t)(date - 1) * grid + g) * 4u;
                const float* cell = (const float*)(table + off);
                lo = cell[0];
                hi = cell[1];
            } else {
                lo = hi = 0.f;
            }
        };
        float x = 0.f;
        int step = 0;
        constexpr int B = 4;
        for (; step + B < last && !(nd && qd); step += B) { // dates step+1 .. step+B < M
            float z[B];
            normal_block_f32<B>(plo, phh, (uint32_t)step, k0, k1, z);
            float s[B], lo[B], hi[B], u[B];
            uint32_t g[B];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                x = x + __builtin_fmaf(sig2, z[j], mu2);
                s[j] = s0 * fast_exp2(x);
                locate(step + j + 1, s[j], lo[j], hi[j], u[j], g[j]);
            }
#pragma unroll
            for (int j = 0; j < B; ++j) visit(step + j + 1, s[j], lo[j], hi[j], u[j], g[j]);
        }
        for (; step < last && !(nd && qd); ++step) {
            const float z = inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1));
            x = x + __builtin_fmaf(sig2, z, mu2);
            const float s = s0 * fast_exp2(x);
            if (step + 1 == last) { // lsm.cpp:76-81: the last date exercises whatever is left
                if (!nd) { nd = true; nds = last; ns = s; }
                if (!qd) { qd = true; qds = last; qs = s; }
            } else {
                float lo, hi, u;
                uint32_t g;
                locate(step + 1, s, lo, hi, u, g);
                visit(step + 1, s, lo, hi, u, g);
            }
        }
    }
    const double cn = nds ? intrinsic_v((double)ns, a.strike, a.type) * disc_pow(a.disc, nds) : 0.0;
    const double cq = qds ? intrinsic_v((double)qs, a.strike, a.type) * disc_pow(a.disc, qds) : 0.0;
    policy_tile_close<MODE != 2, MODE != 0>(a, q, tile, t, cn, cq, dec, dis, fb);
}

// ---- American Greeks by frozen-policy replay ------------------------------------------------
// Central differences of frozen-policy prices (reprice_against_frozen_policy,
// lsm.cpp:151-213, as the reference's tests/lsm_test.cpp:138-170 forms its delta): the QR
// policy fitted at the base parameters exercises five scenarios of each path -- base, spot
// up/down, vol up/down -- generated from the same draws in one pass, and the per-path
// differences accumulate like any Monte Carlo sample.
struct AmGreekArgs {
    double spot_up, spot_dn;             // FP64 spot scenarios
    double drift_u, vsd_u, drift_d, vsd_d; // FP64 vol scenarios (per-date constants)
    float s0_up, s0_dn;                  // FP32 spot scenarios (shared log2 path)
    float mu2_u, sig2_u, mu2_d, sig2_d;  // FP32 vol scenarios
    double inv_2hs, inv_hs2, inv_2hv;
    Welford* parts;                      // [4][tiles + 1]: price, delta, gamma, vega
    uint64_t tiles;
};

// One scenario's stop: whether it has exercised, the date and the exercise value.
struct Stop {
    bool done;
    int date;
    double value;
};

// One scenario's exercise by the frozen policy at `date` (lsm.cpp:110-122).
__device__ __forceinline__ void frozen_visit(const LsmArgs& a, const uint8_t* has, Stop& st,
                                             int date, double s) {
    if (st.done) return;
    const double ex = intrinsic_v(s, a.strike, a.type);
    if (date == a.m) { // lsm.cpp:76-81
        st = Stop{true, date, ex};
        return;
    }
    if (ex > 0.0 && has[date]) {
        double b[4];
        laguerre4(s * a.inv_strike, b);
        const double* beta = a.coeffs + 4 * date;
        double cont = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) cont += beta[j] * b[j];
        if (ex > cont) st = Stop{true, date, ex};
    }
}

// Cash flow of a stopped scenario discounted to t = 0 as backward induction discounts it.
__device__ __forceinline__ double discounted(const LsmArgs& a, const Stop& st) {
    return st.value * disc_pow(a.disc, st.date);
}

// One path per thread; a block of 1024 threads is one canonical tile, partials rebuilt in
// the canonical order as policy_tile_close does.
template <bool FP64>
__global__ __launch_bounds__(4 * kTileLanes) void lsm_am_greeks_kernel(LsmArgs a, AmGreekArgs g) {
    __shared__ double sh[4][kTilePaths];
    __shared__ Welford lds[4][4];
    extern __shared__ uint8_t sh_has[]; // [m + 1], launched with that much dynamic LDS
    const uint64_t tile = blockIdx.x;
    const uint32_t t = threadIdx.x;
    for (int i = (int)t; i <= a.m; i += (int)blockDim.x) sh_has[i] = a.has[i];
    __syncthreads();
    const uint64_t p = tile * kTilePaths + t;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    Stop c0 = {false, 0, 0.0}, c1 = c0, c2 = c0, c3 = c0, c4 = c0;
    auto all_done = [&] { return c0.done && c1.done && c2.done && c3.done && c4.done; };
    if (p < a.n_paths) {
        // t < 1024 = kTilePaths: the block's paths share p >> 32, a scalar here.
        const uint32_t plo = (uint32_t)p, phh = (uint32_t)((tile * kTilePaths) >> 32);
        if constexpr (FP64) {
            double s0 = a.spot, s1 = g.spot_up, s2 = g.spot_dn, s3 = a.spot, s4 = a.spot;
            for (int step = 0; step < a.m && !all_done(); ++step) {
                const double z = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
                const double f = exp(a.drift + a.vsd * z); // the store kernel's update per scenario
                s0 = s0 * f;
                s1 = s1 * f;
                s2 = s2 * f;
                s3 = s3 * exp(g.drift_u + g.vsd_u * z);
                s4 = s4 * exp(g.drift_d + g.vsd_d * z);
                const int d = step + 1;
                frozen_visit(a, sh_has, c0, d, s0);
                frozen_visit(a, sh_has, c1, d, s1);
                frozen_visit(a, sh_has, c2, d, s2);
                frozen_visit(a, sh_has, c3, d, s3);
                frozen_visit(a, sh_has, c4, d, s4);
            }
        } else {
            // frozen_visit's decisions, screened: the date's continuation cubic is formed once
            // for the five scenarios, each scenario decides by decide_screen's estimate, and
            // only the lanes it leaves unsure take the exact FP64 decision (the same result)
            float x = 0.f, xu = 0.f, xd = 0.f;
            const float s0 = (float)a.spot;
            for (int step = 0; step < a.m && !all_done(); ++step) {
                const float z = inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1));
                x = x + __builtin_fmaf(a.sig2, z, a.mu2);
                xu = xu + __builtin_fmaf(g.sig2_u, z, g.mu2_u);
                xd = xd + __builtin_fmaf(g.sig2_d, z, g.mu2_d);
                const float e = fast_exp2(x);
                const int d = step + 1;
                const double sc[5] = {(double)(s0 * e), (double)(g.s0_up * e), (double)(g.s0_dn * e),
                                      (double)(s0 * fast_exp2(xu)), (double)(s0 * fast_exp2(xd))};
                Stop* st[5] = {&c0, &c1, &c2, &c3, &c4};
                if (d == a.m) { // lsm.cpp:76-81
#pragma unroll
                    for (int i = 0; i < 5; ++i)
                        if (!st[i]->done) *st[i] = Stop{true, d, intrinsic_v(sc[i], a.strike, a.type)};
                } else if (sh_has[d]) { // uniform
                    const double* beta = a.coeffs + 4 * d;
                    const ContPoly cp = cont_poly(beta);
                    bool ex[5], unsure[5], any = false;
#pragma unroll
                    for (int i = 0; i < 5; ++i) {
                        unsure[i] = !st[i]->done && decide_screen(a, sc[i], cp, &ex[i]);
                        any = any || unsure[i];
                    }
                    if (any) {
#pragma unroll
                        for (int i = 0; i < 5; ++i)
                            if (unsure[i]) ex[i] = decide(a, sc[i], beta);
                    }
#pragma unroll
                    for (int i = 0; i < 5; ++i)
                        if (!st[i]->done && ex[i]) *st[i] = Stop{true, d, intrinsic_v(sc[i], a.strike, a.type)};
                }
            }
        }
    }
    const double v[5] = {discounted(a, c0), discounted(a, c1), discounted(a, c2), discounted(a, c3),
                         discounted(a, c4)};
    sh[0][t] = v[0];
    sh[1][t] = (v[1] - v[2]) * g.inv_2hs;
    sh[2][t] = ((v[1] - 2.0 * v[0]) + v[2]) * g.inv_hs2;
    sh[3][t] = (v[3] - v[4]) * g.inv_2hv;
    __syncthreads();
    Welford w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = Welford{0, 0.0, 0.0};
    if (t < (uint32_t)kTileLanes) {
#pragma unroll
        for (int k = 0; k < kPathsPerLane; ++k) {
            const uint32_t i = (uint32_t)(k * kTileLanes) + t;
            if (tile * kTilePaths + i >= a.n_paths) break;
            const double rk = 1.0 / (double)(k + 1);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double xv = sh[q][i];
                w[q].n += 1;
                const double dd = xv - w[q].mean;
                w[q].mean += dd * rk;
                w[q].m2 += dd * (xv - w[q].mean);
            }
        }
    }
    const int lane = t & 63;
    if (t < (uint32_t)kTileLanes) { // the 256-lane tree in waves 0..3 (wave-uniform)
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                Welford o;
                o.n = __shfl_down(w[q].n, off, 64);
                o.mean = __shfl_down(w[q].mean, off, 64);
                o.m2 = __shfl_down(w[q].m2, off, 64);
                if ((lane & (2 * off - 1)) == 0) w[q] = welford_merge(w[q], o);
            }
        }
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) lds[q][t >> 6] = w[q];
    }
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            g.parts[(uint64_t)q * (g.tiles + 1) + tile] =
                welford_merge(welford_merge(lds[q][0], lds[q][1]), welford_merge(lds[q][2], lds[q][3]));
}

namespace {

struct LsmRun {
    LsmArgs a;
    uint64_t blocks, tiles;
    uint32_t runs; // independent runs in the batch (grid y)
};

// Lays out the scratch: [store][values][discount powers][Gram sums][date preconditioners]
// [normal-equation sums][tile partials][coeffs][has] followed by `extra_bytes` for the caller
// (returned in *extra when non-null). `with_store` false leaves out the path store, the values
// and the regression's buffers (streaming pricing). `runs` > 1 lays out that
// many independent runs, each region holding the runs one after another (the tile partials:
// runs x tiles, then one merged result per run).
int lsm_setup(double spot, double strike, double rate, double carry, double vol, double time,
              uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision, ThreadCtx* ctx,
              LsmRun* run, size_t extra_bytes = 0, char** extra = nullptr,
              bool with_store = true, uint32_t runs = 1) {
    if (type > 1 || m < 1 || n == 0 || n > (1ull << 30)) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    const bool fp64 = precision == MCDX_PREC_FP64;
    LsmArgs a{};
    const double dt = time / (double)m;
    a.spot = spot;
    a.strike = strike;
    a.inv_strike = 1.0 / strike;
    a.dt = dt;
    a.disc = exp(-rate * dt);
    a.disc_m = disc_pow(a.disc, m);
    a.drift = (rate - carry - 0.5 * vol * vol) * dt;
    a.vsd = vol * sqrt(dt);
    const double log2e = 1.44269504088896340735992468100189214;
    a.mu2 = (float)(a.drift * log2e);
    a.sig2 = (float)(a.vsd * log2e);
    a.seed = seed;
    a.n_paths = n;
    a.m = m;
    a.type = type;
    const uint64_t blocks = (n + kLsmPathLanes - 1) / kLsmPathLanes;
    const uint64_t tiles = tiles_for(n);
    uint64_t bb = (n + 4095) / 4096;
    if (bb > kLsmMaxBlocks) bb = kLsmMaxBlocks;
    a.blocks = (uint32_t)bb;
    a.per_lane = (uint32_t)((n + bb * kLsmLanes - 1) / (bb * kLsmLanes));
    a.sample_n = n < kLsmSample ? n : kLsmSample;
    if (runs < 1 || (runs > 1 && !with_store)) return MCDX_EINVAL;
    const size_t store_bytes = with_store ? (size_t)m * n * (fp64 ? 8 : 4) : 0;
    a.run_store_bytes = runs > 1 ? store_bytes : 0;
    a.run_cf = runs > 1 ? n : 0;
    const uint64_t gram_count = (uint64_t)(m > 1 ? m - 1 : 1) * kGram * kLsmGramBlocks;
    a.run_gram = runs > 1 ? gram_count : 0;
    a.run_ne = runs > 1 ? 2 * (uint64_t)kNe * kLsmMaxBlocks : 0;
    a.run_tiles = runs > 1 ? tiles : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    // The tile partials (whose tail holds the merged results), the coefficients and the flags
    // are adjacent, so one copy brings a single run's results and policy back.
    const size_t o_store = take(store_bytes * runs), o_val = take(with_store ? n * 8 * runs : 0),
                 o_dpow = take(((size_t)m + 1) * 8),
                 o_gram = take(with_store ? gram_count * 8 * runs : 0),
                 o_dfac = take(with_store ? ((size_t)m + 1) * sizeof(DateFactor) * runs : 0),
                 o_ne = take(with_store ? 2 * (size_t)kNe * kLsmMaxBlocks * 8 * runs : 0),
                 o_part = take((tiles + 1) * runs * sizeof(Welford)),
                 o_coeffs = take((size_t)(m + 1) * 4 * 8 * runs),
                 o_has = take(((size_t)m + 1) * runs),
                 o_sync = take(with_store ? (size_t)kLsmSyncStride * 4 * runs : 0),
                 o_extra = take(extra_bytes);
    int st = ensure_dev(ctx, off);
    if (st) return st;
    char* base = (char*)ctx->dev;
    if (extra) *extra = base + o_extra;
    a.store = base + o_store;
    a.val = (double*)(base + o_val);
    a.dpow = (double*)(base + o_dpow);
    a.coeffs = (double*)(base + o_coeffs);
    a.has = (uint8_t*)(base + o_has);
    a.gram = (double*)(base + o_gram);
    a.factors = (DateFactor*)(base + o_dfac);
    a.ne = (double*)(base + o_ne);
    a.partials = (Welford*)(base + o_part);
    a.sync = (uint32_t*)(base + o_sync);
    run->a = a;
    run->blocks = blocks;
    run->tiles = tiles;
    run->runs = runs;
    return MCDX_OK;
}

template <bool FP64>
hipError_t enqueue_paths(const LsmRun& r, hipStream_t s) {
    if (r.runs != 1) return hipErrorInvalidValue; // Philox paths: one run at a time
    const unsigned g = (unsigned)r.blocks;
    hipLaunchKernelGGL(lsm_paths_kernel<FP64>, dim3(g), dim3(256), 0, s, r.a);
    return launch_check();
}

// Backward-induction schedule (mcdx_lsm_schedule): 0 automatic, 1 one launch per date.
std::atomic<int> g_lsm_schedule{MCDX_LSM_AUTO};

// Completion of the last resident backward induction launched on each device.
constexpr int kLsmMaxDevices = 64;
std::mutex g_resident_mu;
hipEvent_t g_resident_done[kLsmMaxDevices] = {};

// Whether one resident launch can run every date of this batch: at most kLsmChunk paths per
// lane, the discount powers within LDS, and every block of the grid resident at once on the
// current device (occupancy per CU x CUs, cached per device).
template <bool FP64, bool CAP>
bool resident_fits(const LsmRun& r) {
    if (g_lsm_schedule.load(std::memory_order_relaxed) != MCDX_LSM_AUTO) return false;
    if (r.a.per_lane > kLsmChunk || r.a.m > kLsmResidentMaxDates) return false;
    static std::atomic<int> capacity[kLsmMaxDevices]; // resident blocks + 1; 0 = not yet known
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kLsmMaxDevices) return false;
    int c = capacity[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        int occ = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, lsm_resident_kernel<FP64, CAP>, kLsmResidentLanes, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            occ = 0;
            cus = 0;
        }
        c = occ * cus + 1;
        capacity[dev].store(c, std::memory_order_relaxed);
    }
    return (uint64_t)r.a.blocks * r.runs <= (uint64_t)(c - 1);
}

// Where the store's paths come from: Philox (the reference's, lsm.cpp:35-58) or Sobol +
// Brownian bridge, optionally digitally shifted (replicate `replicate` of key shift_seed).
struct PathSource {
    bool qmc = false;
    bool shifted = false;
    uint64_t shift_seed = 0;
    uint32_t replicate = 0;
};

// Backward induction with regression (fit) or with a given policy already in coeffs/has.
template <bool FP64>
hipError_t enqueue_backward(const LsmRun& r, bool fit, hipStream_t s) {
    const dim3 g(r.a.blocks, r.runs);
    const bool cap = fit && r.a.cap_y;
    const bool resident = cap ? resident_fits<FP64, true>(r) : resident_fits<FP64, false>(r);
    if (fit && r.a.m > 1) { // every date's preconditioner, before the dates' kernels
        // gb <= 32 blocks per date, at least 16 sample paths per lane
        uint64_t b = (r.a.sample_n + 16 * kLsmLanes - 1) / (16 * kLsmLanes);
        if (b > kLsmGramBlocks) b = kLsmGramBlocks;
        const uint32_t gb = (uint32_t)b;
        const uint32_t per = (uint32_t)((r.a.sample_n + b * kLsmLanes - 1) / (b * kLsmLanes));
        hipLaunchKernelGGL(lsm_gram_kernel<FP64>, dim3(gb, r.a.m - 1, r.runs), dim3(kLsmLanes), 0,
                           s, r.a, gb, per);
        hipLaunchKernelGGL(lsm_gram_merge_kernel<FP64>, dim3(r.a.m - 1, r.runs), dim3(kLsmLanes), 0,
                           s, r.a, gb);
    }
    if (resident) {
        // One resident grid at a time on the device: two that shared it could each hold CUs the
        // other's waiting blocks need. Each launch waits (on the device) for the previous one,
        // on whichever stream of this process it ran.
        int dev = 0;
        hipError_t ge = hipGetDevice(&dev);
        if (ge != hipSuccess) return ge;
        std::lock_guard<std::mutex> lock(g_resident_mu);
        hipEvent_t& last = g_resident_done[dev];
        if (last) ge = hipStreamWaitEvent(s, last, 0);
        else ge = hipEventCreateWithFlags(&last, hipEventDisableTiming);
        if (ge != hipSuccess) return ge;
        if (cap)
            hipLaunchKernelGGL((lsm_resident_kernel<FP64, true>), g, dim3(kLsmResidentLanes), 0, s, r.a, 1);
        else
            hipLaunchKernelGGL((lsm_resident_kernel<FP64, false>), g, dim3(kLsmResidentLanes), 0, s, r.a,
                               fit ? 1 : 0);
        ge = hipEventRecord(last, s);
        if (ge != hipSuccess) return ge;
    } else {
        hipLaunchKernelGGL(lsm_init_cf_kernel<FP64>, dim3((unsigned)r.blocks, r.runs), dim3(256), 0, s,
                           r.a, fit ? 1 : 0);
        for (int date = r.a.m - 1; date >= 1; --date) {
            if (cap)
                hipLaunchKernelGGL((lsm_date_kernel<FP64, true>), g, dim3(kLsmLanes), 0, s, r.a, date, 1, 1);
            else
                hipLaunchKernelGGL((lsm_date_kernel<FP64, false>), g, dim3(kLsmLanes), 0, s, r.a,
                                   date, fit ? 1 : 0, fit ? 1 : 0);
        }
        // date 0: the decisions at date 1 (fitting date 1 first when fitting), no regression
        hipLaunchKernelGGL((lsm_date_kernel<FP64, false>), g, dim3(kLsmLanes), 0, s, r.a, 0, 0,
                           fit ? 1 : 0);
    }
    hipLaunchKernelGGL(lsm_welford_kernel<FP64>, dim3((unsigned)r.tiles, r.runs), dim3(kTileLanes), 0,
                       s, r.a);
    hipError_t e = launch_check();
    if (e != hipSuccess) return e;
    // run k's result at partials[runs x tiles + k]
    return launch_final_merge(r.a.partials, nullptr, r.tiles, r.runs,
                              r.a.partials + r.tiles * r.runs, s);
}

} // namespace

// QR-LSM over `runs` stores: Philox paths (runs = 1), or Sobol + bridge paths, shifted by
// replicates src.replicate .. src.replicate + runs - 1 when src.shifted; the runs share every
// launch (grid y = run). out[k] = run k's price after the inception rule; the policy and the
// inception flag are those of run 0.
int lsm_price_src(double spot, double strike, double rate, double carry, double vol, double time,
                  uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision,
                  const PathSource& src, uint32_t runs, mcdx_result* out, double* coeffs_out,
                  uint8_t* has_out, uint8_t* inception_out) {
    if (!out || runs < 1 || (runs > 1 && !src.qmc)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    LsmRun r;
    char* scratch = nullptr;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r,
                        src.qmc ? qmc_store_scratch_bytes(m, runs) : 0,
                        src.qmc ? &scratch : nullptr, true, runs)))
        return st;
    // One run: its merged result, coefficients and flags come back in one copy of the adjacent
    // regions (lsm_setup); a batch: the runs' results only.
    const char* res_dev = (const char*)(r.a.partials + r.tiles * runs);
    const size_t span = runs == 1 ? (size_t)((const char*)(r.a.has + (m + 1)) - res_dev)
                                  : (size_t)runs * sizeof(Welford);
    if ((st = ensure_pinned(ctx, span))) return st;
    const bool fp64 = precision == MCDX_PREC_FP64;
    hipError_t e;
    if (src.qmc) {
        if ((st = qmc_fill_store(spot, rate, carry, vol, time, m, n, fp64, src.shifted,
                                 src.shift_seed, src.replicate, runs, r.a.store, scratch,
                                 ctx->stream)))
            return st;
        e = hipSuccess;
    } else {
        e = fp64 ? enqueue_paths<true>(r, ctx->stream) : enqueue_paths<false>(r, ctx->stream);
    }
    if (e == hipSuccess)
        e = fp64 ? enqueue_backward<true>(r, true, ctx->stream) : enqueue_backward<false>(r, true, ctx->stream);
    char* h = (char*)ctx->pinned;
    if (e == hipSuccess) e = hipMemcpyAsync(h, res_dev, span, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    if (runs == 1) {
        const double* hc = (const double*)(h + ((const char*)r.a.coeffs - res_dev));
        const uint8_t* hh = (const uint8_t*)(h + ((const char*)r.a.has - res_dev));
        if (coeffs_out)
            for (int i = 0; i < (m + 1) * 4; ++i) coeffs_out[i] = hc[i];
        if (has_out)
            for (int i = 0; i <= m; ++i) has_out[i] = hh[i];
    }
    // inception decision (lsm.cpp:141-148)
    const double inc = intrinsic_v(spot, strike, type);
    for (uint32_t k = 0; k < runs; ++k) // a resident backward induction that gave up waiting
        if (!(((const Welford*)h)[k].mean == ((const Welford*)h)[k].mean)) return MCDX_EDEVICE;
    for (uint32_t k = 0; k < runs; ++k) {
        const mcdx_result res = finish(((const Welford*)h)[k]);
        const bool exercise = inc > res.price;
        if (k == 0 && inception_out) *inception_out = exercise ? 1 : 0;
        out[k] = exercise ? mcdx_result{inc, 0.0, res.path_count} : res;
    }
    return MCDX_OK;
}

int lsm_price(double spot, double strike, double rate, double carry, double vol, double time,
              uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision, mcdx_result* out,
              double* coeffs_out, uint8_t* has_out, uint8_t* inception_out) {
    return lsm_price_src(spot, strike, rate, carry, vol, time, type, m, n, seed, precision,
                         PathSource{}, 1, out, coeffs_out, has_out, inception_out);
}

// Device bytes of path stores one batch of shifted QMC-LSM runs may hold.
constexpr size_t kLsmBatchStoreBytes = (size_t)2 << 30;

// QR-LSM on Sobol + bridge paths: one plain run (replicates == 0) or `replicates` digitally
// shifted runs whose prices give the estimate and its standard error.
int lsm_qmc(double spot, double strike, double rate, double carry, double vol, double time,
            uint8_t type, int32_t m, uint64_t n, uint32_t replicates, uint64_t shift_seed,
            int precision, mcdx_result* out, double* replicate_prices) {
    if (!out || replicates == 1 || replicates > 4096 || n > (1ull << 32) - 2) return MCDX_EINVAL;
    if (replicates == 0)
        return lsm_price_src(spot, strike, rate, carry, vol, time, type, m, n, 0, precision,
                             PathSource{true, false, 0, 0}, 1, out, nullptr, nullptr, nullptr);
    // The replicates run as batches sharing every launch, as many per batch as keep the
    // batch's stores within kLsmBatchStoreBytes.
    const size_t store_bytes = (size_t)m * n * (precision == MCDX_PREC_FP64 ? 8 : 4);
    uint64_t per = store_bytes > 0 ? kLsmBatchStoreBytes / store_bytes : replicates;
    per = per < 1 ? 1 : (per > replicates ? replicates : per);
    double sum = 0.0, ss = 0.0;
    std::vector<double> prices(replicates);
    std::vector<mcdx_result> batch(per);
    for (uint32_t r0 = 0; r0 < replicates; r0 += (uint32_t)per) {
        const uint32_t k = replicates - r0 < per ? replicates - r0 : (uint32_t)per;
        const int st = lsm_price_src(spot, strike, rate, carry, vol, time, type, m, n, 0, precision,
                                     PathSource{true, true, shift_seed, r0}, k, batch.data(),
                                     nullptr, nullptr, nullptr);
        if (st) return st;
        for (uint32_t j = 0; j < k; ++j) {
            prices[r0 + j] = batch[j].price;
            sum += batch[j].price;
        }
    }
    const double mean = sum / (double)replicates;
    for (uint32_t r = 0; r < replicates; ++r) ss += (prices[r] - mean) * (prices[r] - mean);
    if (replicate_prices)
        for (uint32_t r = 0; r < replicates; ++r) replicate_prices[r] = prices[r];
    *out = mcdx_result{mean, sqrt(ss / (double)(replicates - 1) / (double)replicates),
                       n * replicates};
    return MCDX_OK;
}

int lsm_reprice(double spot, double strike, double rate, double carry, double vol, double time,
                uint8_t type, int32_t m, uint64_t n, uint64_t seed, const double* coeffs,
                const uint8_t* has, uint8_t inception, int precision, mcdx_result* out) {
    if (!out || !coeffs || !has) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    LsmRun r;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r)))
        return st;
    const size_t pol_bytes = (size_t)(m + 1) * 4 * 8;
    if ((st = ensure_pinned(ctx, 64 + pol_bytes + (size_t)m + 1))) return st;
    const bool fp64 = precision == MCDX_PREC_FP64;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    char* h = (char*)ctx->pinned;
    for (int i = 0; i < (m + 1) * 4; ++i) ((double*)(h + 64))[i] = coeffs[i];
    for (int i = 0; i <= m; ++i) (h + 64 + pol_bytes)[i] = has[i] ? 1 : 0;
    if (e == hipSuccess) e = fp64 ? enqueue_paths<true>(r, ctx->stream) : enqueue_paths<false>(r, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r.a.coeffs, h + 64, pol_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r.a.has, h + 64 + pol_bytes, (size_t)m + 1, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = fp64 ? enqueue_backward<true>(r, false, ctx->stream) : enqueue_backward<false>(r, false, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream); // pinned staging consumed
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, r.a.partials + r.tiles, sizeof(Welford), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    if (!(((const Welford*)h)->mean == ((const Welford*)h)->mean)) return MCDX_EDEVICE;
    const mcdx_result res = finish(*(const Welford*)h);
    if (inception) // lsm.cpp:207-211
        *out = mcdx_result{intrinsic_v(spot, strike, type), 0.0, res.path_count};
    else
        *out = res;
    return MCDX_OK;
}

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Table range in S/K: from the money boundary to 4.5 standard deviations of log(S_T/K) into
// the money (put: below 1, call: above 1), widened when the spot sits far out of the money.
void table_range(double spot, double strike, double rate, double carry, double vol, double time,
                 uint8_t type, double* lo, double* hi) {
    const double x0 = spot / strike, drift = (rate - carry - 0.5 * vol * vol) * time;
    const double w = 4.5 * vol * sqrt(time);
    if (type == 1) {
        *hi = 1.0;
        *lo = x0 * exp((drift < 0.0 ? drift : 0.0) - w);
        if (!(*lo < 0.95)) *lo = 0.5;
    } else {
        *lo = 1.0;
        *hi = x0 * exp((drift > 0.0 ? drift : 0.0) + w);
        if (!(*hi > 1.05)) *hi = 2.0;
    }
}

// The replay kernels keep the coefficient flags [0..m] in LDS.
constexpr int kMaxLdsDates = 32768;
size_t has_lds_bytes(int m) { return ((size_t)m + 1 + 15) & ~(size_t)15; }

struct PolicyBufs {
    float* table;
    Welford *pn, *pq; // [tiles] partials
    PolicyResults* res;
    uint32_t* tile_counts;
};

size_t policy_bytes(int32_t m, uint32_t grid, uint64_t tiles) {
    return align256((size_t)(m - 1) * grid * 4) + 2 * align256(tiles * sizeof(Welford)) +
           align256(sizeof(PolicyResults)) + align256(3 * tiles * 4);
}

PolicyBufs carve_policy(char* p, int32_t m, uint32_t grid, uint64_t tiles) {
    PolicyBufs b;
    const size_t tb = align256((size_t)(m - 1) * grid * 4), wb = align256(tiles * sizeof(Welford));
    b.table = (float*)p;
    b.pn = (Welford*)(p + tb);
    b.pq = (Welford*)(p + tb + wb);
    b.res = (PolicyResults*)(p + tb + 2 * wb);
    b.tile_counts = (uint32_t*)(p + tb + 2 * wb + align256(sizeof(PolicyResults)));
    return b;
}

// Table (MFMA evaluation of the network on the grid), then the streaming policy kernel, the
// counts and the fixed-order merges of the policies formed (the network's unless mode 2, the
// QR policy's unless mode 0). Results land in *b.res.
// mode: 0 the network (QR outside the table range), 1 both policies (the judge), 2 the QR
// policy alone (no table).
int enqueue_policy(const LsmRun& r, const PolicyBufs& b, const float* w, uint32_t grid,
                   double lo, double hi, bool fp64, int mode, hipStream_t s) {
    const int m = r.a.m;
    if (m + 1 > kMaxLdsDates) return MCDX_EINVAL;
    const uint64_t pts = (uint64_t)(m - 1) * grid;
    const double h = (hi - lo) / (double)(grid - 1);
    if (mode != 2) { // features (4 (x_g - 1), 2 d/M - 1), d = 1..M-1, g < grid
        const MlpEvalGrid eg{grid, m, lo, h, kFeatScale};
        int st = mlp_eval_device(w, nullptr, nullptr, pts, b.table, s, &eg);
        if (st) return st;
    }
    PolicyArgs q{b.table, grid, lo, hi, 1.0 / h, 1.0 / r.a.strike, r.a.strike / kTargetScale,
                 b.pn, b.pq, b.tile_counts, mode};
    // FP32 decisions: the strike rounded outward (up for a put, down for a call), so that
    // phi s > kedge holds exactly when phi (s - K) > 0 for every FP32 spot s.
    const double k = r.a.strike;
    const bool put = r.a.type == 1;
    float kf = (float)k;
    if (put && (double)kf < k) kf = nextafterf(kf, INFINITY);
    if (!put && (double)kf > k) kf = nextafterf(kf, -INFINITY);
    q.kedge = put ? -kf : kf;
    q.phik = (float)(put ? k : -k);
    q.cont_scale_f = (float)(k / kTargetScale);
    q.ua = (float)(1.0 / (k * h));
    q.ub = (float)(-lo / h);
    q.umax = (float)(grid - 1);
    q.gmax = (float)(grid - 2);
    const dim3 g((unsigned)r.tiles), blk(kTilePaths);
    if (fp64)
        hipLaunchKernelGGL(lsm_policy_f64_kernel, g, blk, has_lds_bytes(m), s, r.a, q);
    else {
        using ServeKernel = void (*)(LsmArgs, PolicyArgs);
        static constexpr ServeKernel serve[3][2] = {
            {lsm_serve_kernel<0, false>, lsm_serve_kernel<0, true>},
            {lsm_serve_kernel<1, false>, lsm_serve_kernel<1, true>},
            {lsm_serve_kernel<2, false>, lsm_serve_kernel<2, true>}};
        hipLaunchKernelGGL(serve[mode][put ? 1 : 0], g, blk, has_lds_bytes(m), s, r.a, q);
    }
    hipError_t e = launch_check();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lsm_policy_close_kernel, dim3(mode == 1 ? 3 : 2), dim3(kMergeGroups), 0,
                           s, b.pn, b.pq, r.tiles, b.tile_counts, mode, b.res);
        e = launch_check();
    }
    return status_of(e);
}

// Pinned read-back layout for the policy results.
struct PolicyHost {
    Welford qr_lsm;
    PolicyResults r;
};

// Inception rule (lsm.cpp:141-148) applied to a forward-policy price.
mcdx_result with_inception(const Welford& w, double spot, double strike, uint8_t type) {
    const mcdx_result r = finish(w);
    const double inc = intrinsic_v(spot, strike, type);
    return inc > r.price ? mcdx_result{inc, 0.0, r.path_count} : r;
}

// Largest gap in S/K, over dates with coefficients, between the network's and the QR
// exercise boundaries located on the table grid.
void boundary_gap(const float* table, const double* coeffs, const uint8_t* has, int32_t m,
                  uint32_t grid, double lo, double hi, double strike, uint8_t type, double* gap,
                  uint32_t* dates) {
    const double h = (hi - lo) / (double)(grid - 1);
    *gap = 0.0;
    *dates = 0;
    for (int d = 1; d <= m - 1; ++d) {
        if (!has[d]) continue;
        double bn = NAN, bq = NAN;
        for (uint32_t g = 0; g < grid; ++g) {
            const double x = lo + h * (double)g;
            const double ex = intrinsic_v(x * strike, strike, type);
            if (!(ex > 0.0)) continue;
            const bool en = ex > (double)table[(size_t)(d - 1) * grid + g] * (strike / kTargetScale);
            double b[4];
            laguerre4(x, b);
            double cont = 0.0;
            for (int j = 0; j < 4; ++j) cont += coeffs[4 * d + j] * b[j];
            const bool eq = ex > cont;
            if (type == 1) { // put: exercise below the boundary; keep the largest such x
                if (en) bn = x;
                if (eq) bq = x;
            } else { // call: exercise above; keep the smallest
                if (en && bn != bn) bn = x;
                if (eq && bq != bq) bq = x;
            }
        }
        if (bn == bn && bq == bq) {
            const double gdiff = fabs(bn - bq);
            if (gdiff > *gap) *gap = gdiff;
            ++*dates;
        }
    }
}

} // namespace

int lsm_am_greeks(double spot, double strike, double rate, double carry, double vol, double time,
                  uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision, double hs,
                  double hv, mcdx_result* out4) {
    if (!out4 || !(spot > 0.0) || !(vol > 0.0)) return MCDX_EINVAL;
    if (!(hs > 0.0)) hs = fmax(spot * 0.01, 1e-4); // default_bump_sizes, finite_difference.cpp:31-38
    if (!(hv > 0.0)) hv = fmax(vol * 0.01, 1e-4);
    if (!(hs < spot) || !(hv < vol) || m + 1 > kMaxLdsDates) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const uint64_t tiles = tiles_for(n);
    char* ex = nullptr;
    LsmRun r;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r,
                        4 * (tiles + 1) * sizeof(Welford), &ex)))
        return st;
    if ((st = ensure_pinned(ctx, 4 * sizeof(Welford) + 64))) return st;
    const bool fp64 = precision == MCDX_PREC_FP64;
    hipStream_t s = ctx->stream;
    hipError_t e = fp64 ? enqueue_paths<true>(r, s) : enqueue_paths<false>(r, s);
    if (e == hipSuccess) e = fp64 ? enqueue_backward<true>(r, true, s) : enqueue_backward<false>(r, true, s);
    if (e != hipSuccess) return status_of(e);
    // Scenario constants exactly as the store kernel derives them for those parameters.
    const double dt = time / (double)m, log2e = 1.44269504088896340735992468100189214;
    AmGreekArgs g{};
    g.spot_up = spot + hs;
    g.spot_dn = spot - hs;
    const double vu = vol + hv, vd = vol - hv;
    g.drift_u = (rate - carry - 0.5 * vu * vu) * dt;
    g.vsd_u = vu * sqrt(dt);
    g.drift_d = (rate - carry - 0.5 * vd * vd) * dt;
    g.vsd_d = vd * sqrt(dt);
    g.s0_up = (float)g.spot_up;
    g.s0_dn = (float)g.spot_dn;
    g.mu2_u = (float)(g.drift_u * log2e);
    g.sig2_u = (float)(g.vsd_u * log2e);
    g.mu2_d = (float)(g.drift_d * log2e);
    g.sig2_d = (float)(g.vsd_d * log2e);
    g.inv_2hs = 1.0 / (2.0 * hs);
    g.inv_hs2 = 1.0 / (hs * hs);
    g.inv_2hv = 1.0 / (2.0 * hv);
    g.parts = (Welford*)ex;
    g.tiles = tiles;
    if (fp64)
        hipLaunchKernelGGL(lsm_am_greeks_kernel<true>, dim3((unsigned)tiles), dim3(kTilePaths),
                           has_lds_bytes(m), s, r.a, g);
    else
        hipLaunchKernelGGL(lsm_am_greeks_kernel<false>, dim3((unsigned)tiles), dim3(kTilePaths),
                           has_lds_bytes(m), s, r.a, g);
    e = launch_check();
    for (int q = 0; q < 4 && e == hipSuccess; ++q) {
        Welford* part = g.parts + (uint64_t)q * (tiles + 1);
        e = launch_final_merge(part, nullptr, tiles, 1, part + tiles, s);
    }
    Welford* h = (Welford*)ctx->pinned;
    for (int q = 0; q < 4 && e == hipSuccess; ++q)
        e = hipMemcpyAsync(h + q, g.parts + (uint64_t)q * (tiles + 1) + tiles, sizeof(Welford),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + 4, r.a.partials + r.tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    const mcdx_result fitted = finish(h[4]);
    const double inc = intrinsic_v(spot, strike, type);
    if (inc > fitted.price) { // frozen inception exercise (lsm.cpp:141-148, :207-211)
        const double up = intrinsic_v(spot + hs, strike, type), dn = intrinsic_v(spot - hs, strike, type);
        out4[0] = mcdx_result{inc, 0.0, fitted.path_count};
        out4[1] = mcdx_result{(up - dn) / (2.0 * hs), 0.0, fitted.path_count};
        out4[2] = mcdx_result{((up - 2.0 * inc) + dn) / (hs * hs), 0.0, fitted.path_count};
        out4[3] = mcdx_result{0.0, 0.0, fitted.path_count};
        return MCDX_OK;
    }
    for (int q = 0; q < 4; ++q) out4[q] = finish(h[q]);
    return MCDX_OK;
}

int lsm_learned(double spot, double strike, double rate, double carry, double vol, double time,
                uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision,
                const mcdx_lsm_learn_options* o, float* weights_out, double* coeffs_out,
                uint8_t* has_out, mcdx_lsm_learned_result* out) {
    if (!out || !o || m < 2 || o->steps < 1 || o->steps > (1 << 22) || o->grid_points < 2 ||
        o->grid_points > 4096 || !(o->learning_rate > 0.f) || !(o->final_learning_rate > 0.f))
        return MCDX_EINVAL;
    if ((uint64_t)(m - 1) * n > (1ull << 34)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const uint32_t grid = o->grid_points;
    const size_t cap_bytes = align256((size_t)(m - 1) * n * 4);
    const size_t mlp_bytes = align256(3 * (size_t)kMlpParams * 4) + align256((size_t)o->steps * 4);
    const uint64_t tiles = tiles_for(n);
    char* ex = nullptr;
    LsmRun r;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r,
                        3 * cap_bytes + mlp_bytes + policy_bytes(m, grid, tiles), &ex)))
        return st;
    r.a.cap_f1 = (float*)ex;
    r.a.cap_f2 = (float*)(ex + cap_bytes);
    r.a.cap_y = (float*)(ex + 2 * cap_bytes);
    float* w = (float*)(ex + 3 * cap_bytes);
    float* losses = (float*)(ex + 3 * cap_bytes + align256(3 * (size_t)kMlpParams * 4));
    const PolicyBufs pb = carve_policy(ex + 3 * cap_bytes + mlp_bytes, m, grid, tiles);
    const size_t tbl = (size_t)(m - 1) * grid;
    const size_t h_tbl = align256(sizeof(PolicyHost)), h_loss = h_tbl + align256(tbl * 4),
                 h_coef = h_loss + align256((size_t)o->steps * 4),
                 h_has = h_coef + align256((size_t)(m + 1) * 32),
                 h_w = h_has + align256((size_t)m + 1), h_end = h_w + align256((size_t)kMlpParams * 4);
    if ((st = ensure_pinned(ctx, h_end))) return st;
    const bool fp64 = precision == MCDX_PREC_FP64;
    hipStream_t s = ctx->stream;
    hipError_t e = fp64 ? enqueue_paths<true>(r, s) : enqueue_paths<false>(r, s);
    if (e == hipSuccess) e = fp64 ? enqueue_backward<true>(r, true, s) : enqueue_backward<false>(r, true, s);
    if (e != hipSuccess) return status_of(e);
    // Train on every captured (path, date) sample; OTM samples carry NaN targets.
    if ((st = mlp_init_device(w, o->init_seed, 0.f, s))) return st;
    const float decay = (float)pow((double)o->final_learning_rate / (double)o->learning_rate,
                                   1.0 / (double)o->steps);
    // the last average_steps iterates (all of them when the run is shorter)
    const int avg_from = o->average_steps > 1
                             ? (o->average_steps >= (uint32_t)o->steps ? 0 : o->steps - (int)o->average_steps)
                             : -1;
    if ((st = mlp_train_device(w, r.a.cap_f1, r.a.cap_f2, r.a.cap_y, (uint64_t)(m - 1) * n,
                               o->steps, o->learning_rate, decay, o->sample_seed, losses, nullptr, s,
                               avg_from)))
        return st;
    double lo, hi;
    table_range(spot, strike, rate, carry, vol, time, type, &lo, &hi);
    if ((st = enqueue_policy(r, pb, w, grid, lo, hi, fp64, 1, s))) return st;
    char* h = (char*)ctx->pinned;
    PolicyHost* ph = (PolicyHost*)h;
    e = hipMemcpyAsync(&ph->qr_lsm, r.a.partials + r.tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&ph->r, pb.res, sizeof(PolicyResults), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + h_tbl, pb.table, tbl * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + h_loss, losses, (size_t)o->steps * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + h_coef, r.a.coeffs, (size_t)(m + 1) * 32, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + h_has, r.a.has, (size_t)m + 1, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h + h_w, w, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    const double* hc = (const double*)(h + h_coef);
    const uint8_t* hh = (const uint8_t*)(h + h_has);
    const float* hl = (const float*)(h + h_loss);
    mcdx_lsm_learned_result res{};
    res.qr_lsm = with_inception(ph->qr_lsm, spot, strike, type);
    res.learned = with_inception(ph->r.nn, spot, strike, type);
    res.qr_forward = with_inception(ph->r.qr, spot, strike, type);
    res.itm_decisions = ph->r.ctr[0];
    res.disagreements = ph->r.ctr[1];
    res.fallbacks = ph->r.ctr[2];
    const int last = o->steps < 100 ? o->steps : 100;
    double ls = 0.0;
    for (int i = o->steps - last; i < o->steps; ++i) ls += (double)hl[i];
    res.final_loss = ls / (double)last / (kTargetScale * kTargetScale);
    boundary_gap((const float*)(h + h_tbl), hc, hh, m, grid, lo, hi, strike, type, &res.boundary_gap,
                 &res.boundary_dates);
    res.table_lo = lo;
    res.table_hi = hi;
    if (coeffs_out)
        for (int i = 0; i < (m + 1) * 4; ++i) coeffs_out[i] = hc[i];
    if (has_out)
        for (int i = 0; i <= m; ++i) has_out[i] = hh[i];
    if (weights_out)
        for (int i = 0; i < kMlpParams; ++i) weights_out[i] = ((const float*)(h + h_w))[i];
    *out = res;
    return MCDX_OK;
}

int lsm_learned_price(double spot, double strike, double rate, double carry, double vol,
                      double time, uint8_t type, int32_t m, uint64_t n, uint64_t seed,
                      int precision, const float* weights, uint32_t grid, double lo, double hi,
                      const double* coeffs, const uint8_t* has, uint8_t judge,
                      mcdx_lsm_learned_result* out) {
    if (!out || (!weights && judge != 2) || !coeffs || !has || m < 2 || grid < 2 || grid > 4096 ||
        !(hi > lo) || judge > 2)
        return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const uint64_t tiles = tiles_for(n);
    char* ex = nullptr;
    LsmRun r;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r,
                        align256((size_t)kMlpParams * 4) + policy_bytes(m, grid, tiles), &ex, false)))
        return st;
    float* w = (float*)ex;
    const PolicyBufs pb = carve_policy(ex + align256((size_t)kMlpParams * 4), m, grid, tiles);
    const size_t tbl = (size_t)(m - 1) * grid;
    // Host staging: results, table, then the policy and the weights laid out as on the device
    // (coefficients, flags and the weights that follow them, lsm_setup), so one copy uploads
    // all three. The result and table areas do not overlap the upload's source, so results
    // can come back without waiting for the upload first.
    const size_t pol_span = (size_t)((const char*)(r.a.has + (m + 1)) - (const char*)r.a.coeffs);
    const size_t h_tbl = align256(sizeof(PolicyHost)), h_coef = h_tbl + align256(tbl * 4),
                 h_has = h_coef + (size_t)((const char*)r.a.has - (const char*)r.a.coeffs),
                 h_w = h_coef + align256(pol_span), h_end = h_w + align256((size_t)kMlpParams * 4);
    if ((st = ensure_pinned(ctx, h_end))) return st;
    hipStream_t s = ctx->stream;
    hipError_t e = hipStreamSynchronize(s); // the pinned staging is free
    if (e != hipSuccess) return status_of(e);
    char* h = (char*)ctx->pinned;
    for (int i = 0; i < (m + 1) * 4; ++i) ((double*)(h + h_coef))[i] = coeffs[i];
    for (int i = 0; i <= m; ++i) ((uint8_t*)(h + h_has))[i] = has[i] ? 1 : 0;
    if (judge != 2)
        for (int i = 0; i < kMlpParams; ++i) ((float*)(h + h_w))[i] = weights[i];
    const bool one_copy = (size_t)((const char*)w - (const char*)r.a.coeffs) == h_w - h_coef;
    if (judge != 2 && one_copy) {
        e = hipMemcpyAsync(r.a.coeffs, h + h_coef, h_w - h_coef + (size_t)kMlpParams * 4,
                           hipMemcpyHostToDevice, s);
    } else {
        e = hipMemcpyAsync(r.a.coeffs, h + h_coef, pol_span, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && judge != 2)
            e = hipMemcpyAsync(w, h + h_w, (size_t)kMlpParams * 4, hipMemcpyHostToDevice, s);
    }
    if (e != hipSuccess) return status_of(e);
    if ((st = enqueue_policy(r, pb, w, grid, lo, hi, precision == MCDX_PREC_FP64, judge, s)))
        return st;
    PolicyHost* ph = (PolicyHost*)h;
    e = hipMemcpyAsync(&ph->r, pb.res, sizeof(PolicyResults), hipMemcpyDeviceToHost, s);
    // The table comes back only for the judge's boundary comparison: serving without the
    // judge skips the read-back and the host scan of (m - 1) x grid points.
    if (e == hipSuccess && judge == 1)
        e = hipMemcpyAsync(h + h_tbl, pb.table, tbl * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    mcdx_lsm_learned_result res{};
    res.qr_lsm = mcdx_result{NAN, NAN, 0};
    res.learned = judge != 2 ? with_inception(ph->r.nn, spot, strike, type) : mcdx_result{NAN, NAN, 0};
    res.qr_forward = judge ? with_inception(ph->r.qr, spot, strike, type) : mcdx_result{NAN, NAN, 0};
    res.itm_decisions = ph->r.ctr[0];
    res.disagreements = ph->r.ctr[1];
    res.fallbacks = ph->r.ctr[2];
    res.final_loss = NAN;
    if (judge == 1) {
        boundary_gap((const float*)(h + h_tbl), coeffs, has, m, grid, lo, hi, strike, type,
                     &res.boundary_gap, &res.boundary_dates);
    } else {
        res.boundary_gap = NAN;
        res.boundary_dates = 0;
    }
    res.table_lo = lo;
    res.table_hi = hi;
    *out = res;
    return MCDX_OK;
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_lsm_schedule(int schedule) {
    if (schedule != MCDX_LSM_AUTO && schedule != MCDX_LSM_PER_DATE) return -MCDX_EINVAL;
    return g_lsm_schedule.exchange(schedule);
}

int mcdx_lsm_american_qmc(double spot, double strike, double rate, double carry_yield, double vol,
                          double time, uint8_t type, int32_t monitoring_points,
                          uint64_t path_count, uint32_t replicates, uint64_t shift_seed,
                          int precision, mcdx_result* out, double* replicate_prices) {
    return lsm_qmc(spot, strike, rate, carry_yield, vol, time, type, monitoring_points, path_count,
                   replicates, shift_seed, precision, out, replicate_prices);
}

int mcdx_lsm_american(double spot, double strike, double rate, double carry_yield, double vol,
                      double time, uint8_t type, int32_t monitoring_points, uint64_t path_count,
                      uint64_t seed, int precision, mcdx_result* out, double* coefficients,
                      uint8_t* has_coefficients, uint8_t* exercise_at_inception) {
    return lsm_price(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                     path_count, seed, precision, out, coefficients, has_coefficients,
                     exercise_at_inception);
}

int mcdx_monte_carlo_lsm_american(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type,
                                  int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out) {
    return lsm_price(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                     path_count, seed, mcdx_get_precision(), out, nullptr, nullptr, nullptr);
}

int mcdx_lsm_reprice_frozen(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, int32_t monitoring_points,
                            uint64_t path_count, uint64_t seed, const double* coefficients,
                            const uint8_t* has_coefficients, uint8_t exercise_at_inception,
                            int precision, mcdx_result* out) {
    return lsm_reprice(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                       path_count, seed, coefficients, has_coefficients, exercise_at_inception,
                       precision, out);
}

int mcdx_lsm_american_greeks(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, int32_t monitoring_points,
                            uint64_t path_count, uint64_t seed, int precision, double spot_bump,
                            double vol_bump, mcdx_result* out4) {
    return lsm_am_greeks(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                         path_count, seed, precision, spot_bump, vol_bump, out4);
}

void mcdx_lsm_learn_default_options(mcdx_lsm_learn_options* o) {
    if (!o) return;
    o->init_seed = 7;
    o->sample_seed = 11;
    o->steps = 6000;
    o->learning_rate = 3e-3f;
    o->final_learning_rate = 1e-4f;
    o->grid_points = 1024;
    o->average_steps = 3000;
    o->reserved = 0;
}

int mcdx_lsm_american_learned(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, int32_t monitoring_points,
                              uint64_t path_count, uint64_t seed, int precision,
                              const mcdx_lsm_learn_options* options, float* weights,
                              double* coefficients, uint8_t* has_coefficients,
                              mcdx_lsm_learned_result* out) {
    return lsm_learned(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                       path_count, seed, precision, options, weights, coefficients,
                       has_coefficients, out);
}

int mcdx_lsm_learned_price(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, int precision,
                           const float* weights, uint32_t grid_points, double table_lo,
                           double table_hi, const double* coefficients,
                           const uint8_t* has_coefficients, uint8_t judge,
                           mcdx_lsm_learned_result* out) {
    return lsm_learned_price(spot, strike, rate, carry_yield, vol, time, type, monitoring_points,
                             path_count, seed, precision, weights, grid_points, table_lo,
                             table_hi, coefficients, has_coefficients, judge, out);
}

} // extern "C"
