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
// FP32 in FP32 mode (2^20 paths x 50 dates = 210 MB), FP64 in FP64 mode; the cash flows are
// one FP64 vector. Each date runs one fused kernel: solve the later date's regression from
// its block factors, apply its exercise decision, discount, and fold the ITM rows
// [basis(S/K) | cf] of each lane into an FP64 QR state (R upper-triangular 4x4 plus Q^T y;
// blocked Householder steps per lane), reduce the block's stacked lane factors to one factor
// by Householder reflections with block-wide sums (tall-skinny QR) and write it. The
// least-squares solution is the reference's; only the rounding differs from its sequential
// Householder sweep.
#include <math.h>

#include <vector>

#include "mlp.hpp"
#include "qmc_store.hpp"
#include "runtime.hpp"

namespace mcdx {

constexpr int kLsmBasis = 4;
constexpr int kLsmBlock = 256;

// R (upper triangle, row-major packed 10 values) | z = Q^T y (4) | count of rows folded.
struct QrState {
    double r[10];
    double z[4];
    uint32_t rows;
};

__host__ __device__ __forceinline__ int ridx(int i, int j) { // i <= j
    return i * kLsmBasis - (i * (i - 1)) / 2 + (j - i);
}

// Fold NR rows into (R, z) by Householder reflections of the stacked (4 + NR) x 5 matrix
// [R z; rows ys] (linalg.cpp:7-57 applied to the lane's block): per column, the reflection
// v = x - alpha e1 with alpha = -sign(x1) |x| and tau = 1 / (alpha (alpha - x1)) maps the
// column below the diagonal to zero and is applied to the columns on its right. About a third
// of the FP64 operations of NR Givens row rotations. A column that is zero from the diagonal down
// is left as it is (tau = 0), as is every all-zero row (out-of-the-money paths).
template <int NR>
__device__ __forceinline__ void householder_fold_rows(QrState& st, const double (&rows)[NR][4],
                                                      const double (&ys)[NR]) {
    constexpr int M = kLsmBasis + NR;
    double A[M][kLsmBasis + 1];
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
        for (int j = 0; j < kLsmBasis; ++j) A[i][j] = j >= i ? st.r[ridx(i, j)] : 0.0;
        A[i][kLsmBasis] = st.z[i];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
#pragma unroll
        for (int j = 0; j < kLsmBasis; ++j) A[kLsmBasis + r][j] = rows[r][j];
        A[kLsmBasis + r][kLsmBasis] = ys[r];
    }
#pragma unroll
    for (int j = 0; j < kLsmBasis; ++j) {
        // Rows j+1 .. kLsmBasis-1 are R's rows, zero in column j (below R's diagonal), so
        // the reflection runs over row j and the NR streamed rows only.
        double norm2 = A[j][j] * A[j][j];
#pragma unroll
        for (int i = kLsmBasis; i < M; ++i) norm2 = __builtin_fma(A[i][j], A[i][j], norm2);
        const double x1 = A[j][j];
        const bool live = norm2 > 0.0;
        // |x| = norm2 / sqrt(norm2) and 1 / (alpha (alpha - x1)) by the hardware reciprocal
        // square root and reciprocal refined by Newton steps (within an ulp or two of the IEEE
        // square root and division, at a fraction of their dependent latency).
        const double nrm = live ? norm2 * fast_rsqrt_f64(norm2) : 0.0;
        const double alpha = x1 > 0.0 ? -nrm : nrm;
        const double tau = live ? fast_rcp_f64(alpha * (alpha - x1)) : 0.0;
        const double v0 = x1 - alpha; // v = (v0, A[kLsmBasis.., j])
#pragma unroll
        for (int k = j + 1; k <= kLsmBasis; ++k) {
            double d = v0 * A[j][k];
#pragma unroll
            for (int i = kLsmBasis; i < M; ++i) d = __builtin_fma(A[i][j], A[i][k], d);
            const double f = tau * d;
            A[j][k] = __builtin_fma(-f, v0, A[j][k]);
#pragma unroll
            for (int i = kLsmBasis; i < M; ++i) A[i][k] = __builtin_fma(-f, A[i][j], A[i][k]);
        }
        A[j][j] = live ? alpha : x1;
#pragma unroll
        for (int i = kLsmBasis; i < M; ++i) A[i][j] = live ? 0.0 : A[i][j];
    }
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
        for (int j = i; j < kLsmBasis; ++j) st.r[ridx(i, j)] = A[i][j];
        st.z[i] = A[i][kLsmBasis];
    }
}

// The reference's basis; x^3 / 6 is formed as x^3 * (1/6) (within an ulp of the division,
// without its dependent FP64 division sequence). Every kernel that fits or replays a policy
// evaluates this same function, so fits and replays agree bit for bit.
__host__ __device__ __forceinline__ void laguerre4(double x, double* out) { // lsm.cpp:21-25
    constexpr double kSixth = 1.0 / 6.0;
    const double e = exp(-x / 2.0);
    const double x2 = x * x;
    out[0] = e;
    out[1] = e * (1.0 - x);
    out[2] = e * (1.0 - 2.0 * x + 0.5 * x2);
    out[3] = e * (1.0 - 3.0 * x + 1.5 * x2 - x2 * x * kSixth);
}

__host__ __device__ __forceinline__ double intrinsic_v(double s, double k, int type) {
    const double phi = type == 0 ? 1.0 : -1.0;
    const double v = phi * (s - k);
    return (v < 0.0) ? 0.0 : v;
}

struct LsmArgs {
    double spot, strike, disc, dt;
    double inv_strike;     // 1/K: the regression basis is taken at S * (1/K) (one multiply
                           // per use instead of a division; every decision uses the same)
    double drift, vsd;     // FP64: per-step GBM constants (gbm.cpp:8-10 with dt)
    float mu2, sig2;       // FP32: log2-space step
    uint64_t seed, n_paths;
    int32_t m;             // exercise dates (monitoring points)
    uint32_t blocks;       // backward-induction layout: blocks and paths per lane
    uint32_t per_lane;
    uint8_t type;
    void* store;           // [date 1..M][path], float or double
    double* cf;            // [path]
    double* coeffs;        // [(M+1) x 4] device
    uint8_t* has;          // [M+1] device
    QrState* qr_partials;  // [2][blocks]: the block factors of even and odd dates
    Welford* partials;     // [tiles]
    // Optional capture of the regression samples for the network, [date 1..M-1][path]:
    // features (4 (S/K - 1), 2 date/M - 1) and target 10 cf/K (NaN when out of the money).
    float* cap_f1;
    float* cap_f2;
    float* cap_y;
    // A batch of independent runs of the same contract on different stores (grid y = run):
    // the distance between consecutive runs' store (bytes), cash flows, block factors and
    // tile partials; coeffs and has advance by (M+1) x 4 and M+1. Zero for a single run.
    uint64_t run_store_bytes, run_cf, run_qr, run_tiles;
};

// The arguments of run blockIdx.y of a batch (itself for a single run). The capture
// buffers are not batched: only single runs capture.
__device__ __forceinline__ void select_run(LsmArgs& a) {
    const uint64_t y = blockIdx.y;
    a.store = (char*)a.store + y * a.run_store_bytes;
    a.cf += y * a.run_cf;
    a.coeffs += y * (uint64_t)(a.m + 1) * 4;
    a.has += y * (uint64_t)(a.m + 1);
    a.qr_partials += y * a.run_qr;
    a.partials += y * a.run_tiles;
}

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

// cf = intrinsic at maturity (lsm.cpp:76-81); block 0 of each run also clears the run's
// policy (no coefficients at any date yet).
template <bool FP64>
__global__ __launch_bounds__(256) void lsm_init_cf_kernel(LsmArgs a) {
    select_run(a);
    if (blockIdx.x == 0) {
        for (int i = (int)threadIdx.x; i <= a.m; i += (int)blockDim.x) {
            a.has[i] = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) a.coeffs[4 * i + j] = 0.0;
        }
    }
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    a.cf[p] = intrinsic_v(store_get<FP64>(a, a.m, p), a.strike, a.type);
}

// ---- Backward induction -------------------------------------------------------------------
// Path layout: B = min(256, ceil(n / 4096)) blocks of 256 lanes; lane t of block b owns
// paths ((j * B + b) * 256 + t) for j < P = ceil(n / (256 B)), coalesced per j. Each date
// runs one kernel (reduce the later date's B block factors to one and back-substitute,
// decision at the later date, discount, fold the ITM rows into the lane's QR state, reduce
// the block's lane states to the block factor). The layout depends on n only and every sum
// runs in a fixed order, so the coefficients are bitwise reproducible.
constexpr uint32_t kLsmMaxBlocks = 256;
static_assert(kLsmMaxBlocks <= (uint32_t)kLsmBlock, "merge_solve_all loads one block factor per lane");
constexpr uint32_t kLsmChunk = 8; // paths per lane whose loads are issued together

__host__ __device__ __forceinline__ uint64_t lane_path(uint32_t j, uint32_t block,
                                                       uint32_t blocks, uint32_t t) {
    return ((uint64_t)j * blocks + block) * kLsmBlock + t;
}

// Exercise decision at a later date with coefficients beta (lsm.cpp:110-122).
__device__ __forceinline__ double decide(const LsmArgs& a, double s1, double cf,
                                         const double* beta) {
    const double ex = intrinsic_v(s1, a.strike, a.type);
    if (ex > 0.0) {
        double b[4];
        laguerre4(s1 * a.inv_strike, b);
        double cont = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) cont += beta[j] * b[j];
        if (ex > cont) return ex;
    }
    return cf;
}

// One path's regression row at `date` (lsm.cpp:88-100): the Laguerre basis of S/K when the
// path is in range and in the money, else zeros; returns whether it is. Captures the
// network's training sample for every path in range.
__device__ __forceinline__ bool fold_row(const LsmArgs& a, int date, uint64_t p, bool in,
                                         double s0, double cf, double* row) {
    const bool itm = in && intrinsic_v(s0, a.strike, a.type) > 0.0;
    double b[4];
    laguerre4(s0 * a.inv_strike, b);
#pragma unroll
    for (int k = 0; k < 4; ++k) row[k] = itm ? b[k] : 0.0;
    if (a.cap_y && in) {
        const uint64_t c = (uint64_t)(date - 1) * a.n_paths + p;
        a.cap_f1[c] = (float)(kFeatScale * (s0 / a.strike - 1.0));
        a.cap_f2[c] = (float)(2.0 * (double)date / (double)a.m - 1.0);
        a.cap_y[c] = itm ? (float)(kTargetScale * cf / a.strike) : __builtin_nanf("");
    }
    return itm;
}

// DPP lane permutation of a double (both halves moved by the same control).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the wave's 64 lanes, bit-identical in every lane. Inside each row of 16 lanes:
// lane ^ 1, lane ^ 2 (quad permutations), then the half-row and row mirrors, each pairing a
// lane with one holding the other group's total, so both lanes of a pair add the same two
// operands; then the four row totals as (r0 + r1) + (r2 + r3) from scalar reads.
__device__ __forceinline__ double wave_sum_f64(double v) {
    v += dpp_f64<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v); // row_half_mirror
    v += dpp_f64<0x140>(v); // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Block-wide sums of NV doubles, bit-identical in every lane: wave sums, then the four wave
// totals as (w0 + w1) + (w2 + w3). The block's lane 0 also publishes NP values through the
// same exchange. Consecutive exchanges alternate between two LDS buffers (buf = 0, 1), so one
// barrier per exchange suffices: a buffer is rewritten only after every lane has passed the
// other buffer's barrier, i.e. finished reading this one.
constexpr int kLsmXchg = 4 * 6 + 5; // doubles per exchange buffer (NV <= 6, NP <= 5)
template <int NV, int NP>
__device__ __forceinline__ void block_sums(double (&v)[NV], double (&pub)[NP], double* lds,
                                           int buf) {
    static_assert(4 * NV + NP <= kLsmXchg, "exchange buffer too small");
    double* x = lds + buf * kLsmXchg;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum_f64(v[i]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) x[w * NV + i] = v[i];
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NP; ++i) x[4 * NV + i] = pub[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = (x[i] + x[NV + i]) + (x[2 * NV + i] + x[3 * NV + i]);
#pragma unroll
    for (int i = 0; i < NP; ++i) pub[i] = x[4 * NV + i];
}
constexpr int kLsmLdsDoubles = 2 * kLsmXchg + 8; // two exchange buffers, the solve's results

// One column step of the block's stacked Householder reduction (block_factor). Every lane's
// rows keep the upper-triangular zero pattern of a factor (row r is zero before column r:
// the reflections of earlier columns add multiples of those zeros), so column J has non-zero
// entries only in rows 0 .. J of each lane; lane 0's rows 0 .. J - 1 are finished and its row
// J is the pivot, so the other lanes' rows 0 .. J form the column below the pivot.
template <int J>
__device__ __forceinline__ void stacked_reflect(double (&A)[kLsmBasis][kLsmBasis + 1],
                                                bool lead, double& rows, double* lds) {
    constexpr int NK = kLsmBasis - J;             // columns J+1 .. kLsmBasis (the last is z)
    constexpr int NV = 1 + NK + (J == 0 ? 1 : 0); // sum of squares, dot products, row count
    double v[NV], pub[1 + NK];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = 0.0;
    if (!lead) {
#pragma unroll
        for (int r = 0; r <= J; ++r) {
            const double x = A[r][J];
            v[0] = __builtin_fma(x, x, v[0]);
#pragma unroll
            for (int k = 0; k < NK; ++k) v[1 + k] = __builtin_fma(x, A[r][J + 1 + k], v[1 + k]);
        }
    }
    if constexpr (J == 0) v[NV - 1] = rows;
#pragma unroll
    for (int k = 0; k <= NK; ++k) pub[k] = lead ? A[J][J + k] : 0.0;
    block_sums(v, pub, lds, J & 1);
    if constexpr (J == 0) rows = v[NV - 1];
    // The reflection v = x - alpha e_J of the stacked column (alpha = -sign(x1) |x|,
    // tau = 1 / (alpha (alpha - x1))), applied to the columns on its right; a column that is
    // zero from the pivot down is left as it is.
    const double x1 = pub[0];
    const double norm2 = __builtin_fma(x1, x1, v[0]);
    const bool live = norm2 > 0.0;
    const double nrm = live ? norm2 * fast_rsqrt_f64(norm2) : 0.0; // as householder_fold_rows
    const double alpha = x1 > 0.0 ? -nrm : nrm;
    const double tau = live ? fast_rcp_f64(alpha * (alpha - x1)) : 0.0;
    const double v0 = x1 - alpha;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const double f = tau * __builtin_fma(v0, pub[1 + k], v[1 + k]);
        if (lead) {
            A[J][J + 1 + k] = __builtin_fma(-f, v0, A[J][J + 1 + k]);
        } else {
#pragma unroll
            for (int r = 0; r <= J; ++r) A[r][J + 1 + k] = __builtin_fma(-f, A[r][J], A[r][J + 1 + k]);
        }
    }
    if (lead) {
        A[J][J] = live ? alpha : x1;
    } else {
#pragma unroll
        for (int r = 0; r <= J; ++r) A[r][J] = live ? 0.0 : A[r][J];
    }
}

// The block's combined factor: the lanes' [R | z] rows stacked in lane order (1024 x 5)
// reduced to one upper-triangular [R | z] by Householder reflections, as linalg.cpp:7-57
// reduces the full design matrix. Column j pivots on lane 0's row j; each column costs one
// block-wide sum of the dot products it needs, in a fixed order, so the result is bitwise
// reproducible. The factor is returned in lane 0 (other lanes: partial garbage).
__device__ __forceinline__ QrState block_factor(const QrState& in, double* lds) {
    double A[kLsmBasis][kLsmBasis + 1];
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
        for (int j = 0; j < kLsmBasis; ++j) A[i][j] = j >= i ? in.r[ridx(i, j)] : 0.0;
        A[i][kLsmBasis] = in.z[i];
    }
    const bool lead = threadIdx.x == 0;
    double rows = (double)in.rows; // exact: at most 2^32
    stacked_reflect<0>(A, lead, rows, lds);
    stacked_reflect<1>(A, lead, rows, lds);
    stacked_reflect<2>(A, lead, rows, lds);
    stacked_reflect<3>(A, lead, rows, lds);
    QrState o;
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) {
#pragma unroll
        for (int j = i; j < kLsmBasis; ++j) o.r[ridx(i, j)] = A[i][j];
        o.z[i] = A[i][kLsmBasis];
    }
    o.rows = (uint32_t)rows;
    return o;
}

// Combine nb <= 256 block factors (lane b holds factor b, the other lanes zeros) and back-substitute
// (linalg.cpp:58-69) when at least four ITM rows were folded (lsm.cpp:102-104). Every lane of
// the 256-lane block receives beta and has.
__device__ __forceinline__ void merge_solve_all(const QrState& acc, double* lds, double* beta,
                                                int* has) {
    const QrState t = block_factor(acc, lds);
    double* res = lds + 2 * kLsmXchg; // not touched by block_sums
    if (threadIdx.x == 0) {
        const int h = t.rows >= (uint32_t)kLsmBasis ? 1 : 0;
        double b[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = kLsmBasis - 1; i >= 0; --i) {
            double sum = t.z[i];
            for (int j = i + 1; j < kLsmBasis; ++j) sum -= t.r[ridx(i, j)] * b[j];
            const double diag = t.r[ridx(i, i)];
            b[i] = diag != 0.0 ? sum / diag : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) res[i] = b[i];
        res[4] = (double)h;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) beta[i] = res[i];
    *has = res[4] != 0.0 ? 1 : 0;
}

// One date of the backward induction. With fit, the later date's regression is solved here:
// every block reduces the later date's block factors (written by the previous launch, in the
// other half of the double-buffered partials) in the same fixed order, so every block holds
// the same coefficients, and block 0 records them; no separate solve launch per date. Without
// fit, the policy in coeffs/has is applied.
template <bool FP64>
__global__ __launch_bounds__(kLsmBlock) void lsm_date_kernel(LsmArgs a, int date, int regress,
                                                             int fit) {
    select_run(a);
    __shared__ double lds[kLsmLdsDoubles];
    const int later = date + 1;
    // Chunks of kLsmChunk paths per lane, every load of a chunk issued before its first use;
    // the first chunk's loads go out before the later date's regression is solved, so their
    // latency overlaps it (issuing each next chunk's loads a chunk early measured slower).
    // Loads are unconditional at in-bounds indices (path 0 past the end, store column 1 for
    // the final date 0), so no load waits inside a branch; unused values are ignored.
    const int col = date >= 1 ? date : 1;
    auto load_chunk = [&](uint32_t j0, double (&cf)[kLsmChunk], double (&s1)[kLsmChunk],
                          double (&s0)[kLsmChunk]) {
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
            const uint64_t q = (j0 + u < a.per_lane && p < a.n_paths) ? p : 0;
            cf[u] = a.cf[q];
            s1[u] = store_get<FP64>(a, later, q);
            s0[u] = store_get<FP64>(a, col, q);
        }
    };
    const bool solve = fit && later <= a.m - 1; // uniform
    QrState acc = {};
    if (solve && threadIdx.x < a.blocks) acc = a.qr_partials[(size_t)(later & 1) * a.blocks + threadIdx.x];
    double cf[kLsmChunk], s1[kLsmChunk], s0[kLsmChunk];
    load_chunk(0, cf, s1, s0);
    double beta[4] = {0.0, 0.0, 0.0, 0.0};
    bool use = false;
    if (later <= a.m - 1) { // uniform
        if (fit) {
            int has = 0;
            merge_solve_all(acc, lds, beta, &has);
            use = has != 0;
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                a.has[later] = (uint8_t)has;
                if (has)
                    for (int j = 0; j < 4; ++j) a.coeffs[4 * later + j] = beta[j];
            }
        } else {
            use = a.has[later] != 0;
            if (use)
                for (int j = 0; j < 4; ++j) beta[j] = a.coeffs[4 * later + j];
        }
    }
    QrState st = {};
    // Rows are folded in ascending path order, as one path at a time would.
    for (uint32_t j0 = 0; j0 < a.per_lane; j0 += kLsmChunk) {
        if (j0 > 0) load_chunk(j0, cf, s1, s0);
#pragma unroll
        for (uint32_t u = 0; u < kLsmChunk; ++u) {
            const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
            if (j0 + u >= a.per_lane || p >= a.n_paths) break;
            double c = cf[u];
            if (use) c = decide(a, s1[u], c, beta);
            c *= a.disc; // lsm.cpp:84-86 (and :125-127 after the last date)
            a.cf[p] = c;
            cf[u] = c;
        }
        if (regress) { // uniform across the grid
            // The chunk's rows (out-of-the-money and past-the-end paths as zero rows) folded
            // into the lane's factor in one blocked Householder step.
            double rows[kLsmChunk][4], ys[kLsmChunk];
#pragma unroll
            for (uint32_t u = 0; u < kLsmChunk; ++u) {
                const uint64_t p = lane_path(j0 + u, blockIdx.x, a.blocks, threadIdx.x);
                const bool in = j0 + u < a.per_lane && p < a.n_paths;
                const bool itm = fold_row(a, date, p, in, s0[u], cf[u], rows[u]);
                ys[u] = itm ? cf[u] : 0.0;
                st.rows += itm ? 1u : 0u;
            }
            householder_fold_rows<kLsmChunk>(st, rows, ys);
        }
    }
    if (!regress) return; // uniform across the grid
    const QrState t = block_factor(st, lds);
    if (threadIdx.x == 0) a.qr_partials[(size_t)(date & 1) * a.blocks + blockIdx.x] = t;
}

// Welford over the final cash flows in the canonical tile order.
__global__ __launch_bounds__(kTileLanes) void lsm_welford_kernel(LsmArgs a) {
    select_run(a);
    __shared__ Welford lds4[4];
    const uint64_t tile = blockIdx.x;
    Welford w = {0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPathsPerLane; ++k) {
        const uint64_t p = tile * kTilePaths + (uint64_t)(k * kTileLanes) + threadIdx.x;
        if (p < a.n_paths) {
            const double x = a.cf[p];
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
};

// counters[c] = sum over tiles of tile_counts[c][tile], c < 3 (one block).
__global__ __launch_bounds__(256) void lsm_count_kernel(const uint32_t* tile_counts,
                                                        uint64_t tiles,
                                                        unsigned long long* counters) {
    __shared__ unsigned long long part[3][4];
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
    if (threadIdx.x < 3)
        counters[threadIdx.x] = part[threadIdx.x][0] + part[threadIdx.x][1] +
                                part[threadIdx.x][2] + part[threadIdx.x][3];
}

// Grid features for the table: (4 (x_g - 1), 2 d/M - 1) for d = 1..M-1, g < grid.
__global__ __launch_bounds__(256) void lsm_grid_kernel(float* f1, float* f2, int m, uint32_t grid,
                                                       double xlo, double h) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)(m - 1) * grid) return;
    const int date = (int)(i / grid) + 1;
    const uint32_t g = (uint32_t)(i % grid);
    f1[i] = (float)(kFeatScale * (xlo + h * (double)g - 1.0));
    f2[i] = (float)(2.0 * (double)date / (double)m - 1.0);
}

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

// One path per thread; a block of 1024 threads is one canonical tile (thread t computes path
// tile * 1024 + t), and the tile's Welford partial is rebuilt from LDS in the canonical
// order (lane l < 256 adds paths l, l+256, l+512, l+768 in turn, then the 256-lane tree), so
// it equals the partial of the 256-lane path kernels. Each path is regenerated exactly as the
// store kernel writes it and stopped at each policy's first exercise date; the cash flow is
// then discounted by one multiplication per date, the order backward induction applies them
// in, so the QR policy reproduces the backward-induction price bit for bit.
template <bool FP64>
__global__ __launch_bounds__(4 * kTileLanes) void lsm_policy_kernel(LsmArgs a, PolicyArgs q) {
    __shared__ double sh_n[kTilePaths], sh_q[kTilePaths];
    __shared__ Welford lds_n[4], lds_q[4];
    __shared__ uint32_t sh_cnt[3][16];
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
        if constexpr (FP64) {
            double sp = a.spot;
            for (int step = 0; step < a.m && !(ps.n_done && ps.q_done); ++step) {
                const double z = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
                sp = sp * exp(a.drift + a.vsd * z);
                policy_visit(a, q, sh_has, ps, step + 1, sp);
            }
        } else {
            float x = 0.f;
            const float s0 = (float)a.spot;
            int step = 0;
            constexpr int B = 4;
            for (; step + B <= a.m && !(ps.n_done && ps.q_done); step += B) {
                float z[B];
                normal_block_f32<B>(plo, phh, (uint32_t)step, k0, k1, z);
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    x = x + __builtin_fmaf(a.sig2, z[j], a.mu2);
                    policy_visit(a, q, sh_has, ps, step + j + 1, (double)(s0 * fast_exp2(x)));
                }
            }
            for (; step < a.m && !(ps.n_done && ps.q_done); ++step) {
                const float z = inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1));
                x = x + __builtin_fmaf(a.sig2, z, a.mu2);
                policy_visit(a, q, sh_has, ps, step + 1, (double)(s0 * fast_exp2(x)));
            }
        }
    }
    double cn = ps.n_val, cq = ps.q_val;
    for (int i = 0; i < ps.n_date; ++i) cn *= a.disc;
    for (int i = 0; i < ps.q_date; ++i) cq *= a.disc;
    sh_n[t] = cn;
    sh_q[t] = cq;
    uint32_t c_dec = ps.dec, c_dis = ps.dis, c_fb = ps.fb;
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
    if (t < 3) { // per-tile counts, summed by lsm_count_kernel (no same-address atomics)
        uint32_t sum = 0;
        for (int w = 0; w < 16; ++w) sum += sh_cnt[t][w];
        q.tile_counts[(uint64_t)t * gridDim.x + tile] = sum;
    }
    Welford wn = {0, 0.0, 0.0}, wq = {0, 0.0, 0.0};
    if (t < (uint32_t)kTileLanes) {
#pragma unroll
        for (int k = 0; k < kPathsPerLane; ++k) {
            const uint32_t i = (uint32_t)(k * kTileLanes) + t;
            if (tile * kTilePaths + i >= a.n_paths) break;
            const double rk = 1.0 / (double)(k + 1);
            const double xn = sh_n[i], xq = sh_q[i];
            wn.n += 1;
            double d = xn - wn.mean;
            wn.mean += d * rk;
            wn.m2 += d * (xn - wn.mean);
            wq.n += 1;
            d = xq - wq.mean;
            wq.mean += d * rk;
            wq.m2 += d * (xq - wq.mean);
        }
    }
    // the 256-lane tree of block_merge_256, run by every wave (waves 4..15 hold zeros)
    const int lane = t & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Welford on, oq;
        on.n = __shfl_down(wn.n, off, 64);
        on.mean = __shfl_down(wn.mean, off, 64);
        on.m2 = __shfl_down(wn.m2, off, 64);
        oq.n = __shfl_down(wq.n, off, 64);
        oq.mean = __shfl_down(wq.mean, off, 64);
        oq.m2 = __shfl_down(wq.m2, off, 64);
        if ((lane & (2 * off - 1)) == 0) {
            wn = welford_merge(wn, on);
            wq = welford_merge(wq, oq);
        }
    }
    if (lane == 0 && t < (uint32_t)kTileLanes) {
        lds_n[t >> 6] = wn;
        lds_q[t >> 6] = wq;
    }
    __syncthreads();
    if (t == 0) {
        q.part_nn[tile] = welford_merge(welford_merge(lds_n[0], lds_n[1]),
                                        welford_merge(lds_n[2], lds_n[3]));
        q.part_qr[tile] = welford_merge(welford_merge(lds_q[0], lds_q[1]),
                                        welford_merge(lds_q[2], lds_q[3]));
    }
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

// Cash flow of a stopped scenario discounted one multiplication per date, in backward
// induction's order.
__device__ __forceinline__ double discounted(const LsmArgs& a, const Stop& st) {
    double c = st.value;
    for (int i = 0; i < st.date; ++i) c *= a.disc;
    return c;
}

// One path per thread; a block of 1024 threads is one canonical tile, partials rebuilt in
// the canonical order as lsm_policy_kernel does.
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
            float x = 0.f, xu = 0.f, xd = 0.f;
            const float s0 = (float)a.spot;
            for (int step = 0; step < a.m && !all_done(); ++step) {
                const float z = inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1));
                x = x + __builtin_fmaf(a.sig2, z, a.mu2);
                xu = xu + __builtin_fmaf(g.sig2_u, z, g.mu2_u);
                xd = xd + __builtin_fmaf(g.sig2_d, z, g.mu2_d);
                const float e = fast_exp2(x);
                const int d = step + 1;
                frozen_visit(a, sh_has, c0, d, (double)(s0 * e));
                frozen_visit(a, sh_has, c1, d, (double)(g.s0_up * e));
                frozen_visit(a, sh_has, c2, d, (double)(g.s0_dn * e));
                frozen_visit(a, sh_has, c3, d, (double)(s0 * fast_exp2(xu)));
                frozen_visit(a, sh_has, c4, d, (double)(s0 * fast_exp2(xd)));
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
    if (lane == 0 && t < (uint32_t)kTileLanes)
#pragma unroll
        for (int q = 0; q < 4; ++q) lds[q][t >> 6] = w[q];
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

// Lays out the scratch: [store][cf][coeffs][has][qr partials][tile partials] followed by
// `extra_bytes` for the caller (returned in *extra when non-null). `with_store` false leaves
// out the path store and the cash-flow vector (streaming pricing). `runs` > 1 lays out that
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
    a.drift = (rate - carry - 0.5 * vol * vol) * dt;
    a.vsd = vol * sqrt(dt);
    const double log2e = 1.44269504088896340735992468100189214;
    a.mu2 = (float)(a.drift * log2e);
    a.sig2 = (float)(a.vsd * log2e);
    a.seed = seed;
    a.n_paths = n;
    a.m = m;
    a.type = type;
    const uint64_t blocks = (n + kLsmBlock - 1) / kLsmBlock;
    const uint64_t tiles = tiles_for(n);
    uint64_t bb = (n + 4095) / 4096;
    if (bb > kLsmMaxBlocks) bb = kLsmMaxBlocks;
    a.blocks = (uint32_t)bb;
    a.per_lane = (uint32_t)((n + bb * kLsmBlock - 1) / (bb * kLsmBlock));
    if (runs < 1 || (runs > 1 && !with_store)) return MCDX_EINVAL;
    const size_t store_bytes = with_store ? (size_t)m * n * (fp64 ? 8 : 4) : 0;
    a.run_store_bytes = runs > 1 ? store_bytes : 0;
    a.run_cf = runs > 1 ? n : 0;
    a.run_qr = runs > 1 ? 2 * (uint64_t)a.blocks : 0;
    a.run_tiles = runs > 1 ? tiles : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    // The tile partials (whose tail holds the merged results), the coefficients and the flags
    // are adjacent, so one copy brings a single run's results and policy back.
    const size_t o_store = take(store_bytes * runs), o_cf = take(with_store ? n * 8 * runs : 0),
                 o_qr = take(with_store ? 2 * (size_t)a.blocks * sizeof(QrState) * runs : 0),
                 o_part = take((tiles + 1) * runs * sizeof(Welford)),
                 o_coeffs = take((size_t)(m + 1) * 4 * 8 * runs),
                 o_has = take(((size_t)m + 1) * runs), o_extra = take(extra_bytes);
    int st = ensure_dev(ctx, off);
    if (st) return st;
    char* base = (char*)ctx->dev;
    if (extra) *extra = base + o_extra;
    a.store = base + o_store;
    a.cf = (double*)(base + o_cf);
    a.coeffs = (double*)(base + o_coeffs);
    a.has = (uint8_t*)(base + o_has);
    a.qr_partials = (QrState*)(base + o_qr);
    a.partials = (Welford*)(base + o_part);
    run->a = a;
    run->blocks = blocks;
    run->tiles = tiles;
    run->runs = runs;
    return MCDX_OK;
}

// Cash flows at maturity and a cleared policy, once the store holds the paths.
template <bool FP64>
hipError_t enqueue_init(const LsmRun& r, hipStream_t s) {
    const unsigned g = (unsigned)r.blocks;
    hipLaunchKernelGGL(lsm_init_cf_kernel<FP64>, dim3(g, r.runs), dim3(256), 0, s, r.a);
    return launch_check();
}

template <bool FP64>
hipError_t enqueue_paths(const LsmRun& r, hipStream_t s) {
    if (r.runs != 1) return hipErrorInvalidValue; // Philox paths: one run at a time
    const unsigned g = (unsigned)r.blocks;
    hipLaunchKernelGGL(lsm_paths_kernel<FP64>, dim3(g), dim3(256), 0, s, r.a);
    hipError_t e = launch_check();
    return e == hipSuccess ? enqueue_init<FP64>(r, s) : e;
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
    for (int date = r.a.m - 1; date >= 1; --date)
        hipLaunchKernelGGL(lsm_date_kernel<FP64>, g, dim3(kLsmBlock), 0, s, r.a, date,
                           fit ? 1 : 0, fit ? 1 : 0);
    // date 0: the decisions at date 1 (fitting date 1 first when fitting), no regression
    hipLaunchKernelGGL(lsm_date_kernel<FP64>, g, dim3(kLsmBlock), 0, s, r.a, 0, 0,
                       fit ? 1 : 0);
    hipLaunchKernelGGL(lsm_welford_kernel, dim3((unsigned)r.tiles, r.runs), dim3(kTileLanes), 0,
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
        e = fp64 ? enqueue_init<true>(r, ctx->stream) : enqueue_init<false>(r, ctx->stream);
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
    float *table, *gx, *gt;
    Welford *pn, *pq;
    unsigned long long* ctr;
    uint32_t* tile_counts;
};

size_t policy_bytes(int32_t m, uint32_t grid, uint64_t tiles) {
    return 3 * align256((size_t)(m - 1) * grid * 4) + 2 * align256((tiles + 1) * sizeof(Welford)) +
           256 + align256(3 * tiles * 4);
}

PolicyBufs carve_policy(char* p, int32_t m, uint32_t grid, uint64_t tiles) {
    PolicyBufs b;
    const size_t tb = align256((size_t)(m - 1) * grid * 4), wb = align256((tiles + 1) * sizeof(Welford));
    b.table = (float*)p;
    b.gx = (float*)(p + tb);
    b.gt = (float*)(p + 2 * tb);
    b.pn = (Welford*)(p + 3 * tb);
    b.pq = (Welford*)(p + 3 * tb + wb);
    b.ctr = (unsigned long long*)(p + 3 * tb + 2 * wb);
    b.tile_counts = (uint32_t*)(p + 3 * tb + 2 * wb + 256);
    return b;
}

// Table (MFMA evaluation of the network on the grid), then the streaming policy kernel and
// the two fixed-order merges. Results land in pn[tiles], pq[tiles] and ctr[0..2].
// mode: 0 the network (QR outside the table range), 1 both policies (the judge), 2 the QR
// policy alone (no table).
int enqueue_policy(const LsmRun& r, const PolicyBufs& b, const float* w, uint32_t grid,
                   double lo, double hi, bool fp64, int mode, hipStream_t s) {
    const int m = r.a.m;
    if (m + 1 > kMaxLdsDates) return MCDX_EINVAL;
    const uint64_t pts = (uint64_t)(m - 1) * grid;
    const double h = (hi - lo) / (double)(grid - 1);
    if (mode != 2) {
        hipLaunchKernelGGL(lsm_grid_kernel, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s,
                           b.gx, b.gt, m, grid, lo, h);
        int st = mlp_eval_device(w, b.gx, b.gt, pts, b.table, s);
        if (st) return st;
    }
    PolicyArgs q{b.table, grid, lo, hi, 1.0 / h, 1.0 / r.a.strike, r.a.strike / kTargetScale,
                 b.pn, b.pq, b.tile_counts, mode};
    if (fp64)
        hipLaunchKernelGGL(lsm_policy_kernel<true>, dim3((unsigned)r.tiles), dim3(kTilePaths),
                           has_lds_bytes(m), s, r.a, q);
    else
        hipLaunchKernelGGL(lsm_policy_kernel<false>, dim3((unsigned)r.tiles), dim3(kTilePaths),
                           has_lds_bytes(m), s, r.a, q);
    hipError_t e = launch_check();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lsm_count_kernel, dim3(1), dim3(256), 0, s, b.tile_counts, r.tiles, b.ctr);
        e = launch_check();
    }
    if (e == hipSuccess) e = launch_final_merge(b.pn, nullptr, r.tiles, 1, b.pn + r.tiles, s);
    if (e == hipSuccess) e = launch_final_merge(b.pq, nullptr, r.tiles, 1, b.pq + r.tiles, s);
    return status_of(e);
}

// Pinned read-back layout for the policy results.
struct PolicyHost {
    Welford qr_lsm, nn, qr;
    unsigned long long ctr[3];
    double loss_sum;
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
    if (e == hipSuccess) e = hipMemcpyAsync(&ph->nn, pb.pn + tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&ph->qr, pb.pq + tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ph->ctr, pb.ctr, 3 * 8, hipMemcpyDeviceToHost, s);
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
    res.learned = with_inception(ph->nn, spot, strike, type);
    res.qr_forward = with_inception(ph->qr, spot, strike, type);
    res.itm_decisions = ph->ctr[0];
    res.disagreements = ph->ctr[1];
    res.fallbacks = ph->ctr[2];
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
    // Host staging: results, table, then the policy laid out as on the device (coefficients
    // and flags adjacent, lsm_setup), so one copy uploads both; then the weights. The result
    // and table areas do not overlap the uploads' sources, so results can come back without
    // waiting for the uploads first.
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
    e = hipMemcpyAsync(r.a.coeffs, h + h_coef, pol_span, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && judge != 2)
        e = hipMemcpyAsync(w, h + h_w, (size_t)kMlpParams * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return status_of(e);
    if ((st = enqueue_policy(r, pb, w, grid, lo, hi, precision == MCDX_PREC_FP64, judge, s)))
        return st;
    PolicyHost* ph = (PolicyHost*)h;
    e = hipMemcpyAsync(&ph->nn, pb.pn + tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&ph->qr, pb.pq + tiles, sizeof(Welford), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ph->ctr, pb.ctr, 3 * 8, hipMemcpyDeviceToHost, s);
    // The table comes back only for the judge's boundary comparison: serving without the
    // judge skips the read-back and the host scan of (m - 1) x grid points.
    if (e == hipSuccess && judge == 1)
        e = hipMemcpyAsync(h + h_tbl, pb.table, tbl * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return status_of(e);
    mcdx_lsm_learned_result res{};
    res.qr_lsm = mcdx_result{NAN, NAN, 0};
    res.learned = judge != 2 ? with_inception(ph->nn, spot, strike, type) : mcdx_result{NAN, NAN, 0};
    res.qr_forward = judge ? with_inception(ph->qr, spot, strike, type) : mcdx_result{NAN, NAN, 0};
    res.itm_decisions = ph->ctr[0];
    res.disagreements = ph->ctr[1];
    res.fallbacks = ph->ctr[2];
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
