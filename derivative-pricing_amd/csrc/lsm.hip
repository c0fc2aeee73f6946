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
// one FP64 vector. Each date runs one fused kernel: apply the later date's exercise decision,
// discount, and fold the ITM rows [basis(S/K) | cf] of each lane into an FP64 Givens QR
// state (R upper-triangular 4x4 plus Q^T y); lane states merge along the block's pairwise
// tree, block states along a fixed tree in the solve kernel (tall-skinny QR), which then
// back-substitutes the coefficients on the device. The least-squares solution is the
// reference's; only the rounding differs from its sequential Householder sweep.
#include <math.h>

#include <vector>

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

// Fold one row (a[0..3], y) into (R, z) with Givens rotations.
__host__ __device__ __forceinline__ void qr_add_row(QrState& st, double a0, double a1, double a2,
                                                    double a3, double y) {
    double a[4] = {a0, a1, a2, a3};
#pragma unroll
    for (int j = 0; j < kLsmBasis; ++j) {
        const double aj = a[j];
        if (aj == 0.0) continue;
        const double rjj = st.r[ridx(j, j)];
        const double h = sqrt(rjj * rjj + aj * aj);
        const double c = rjj / h, s = aj / h;
#pragma unroll
        for (int k = j; k < kLsmBasis; ++k) {
            const double rk = st.r[ridx(j, k)];
            st.r[ridx(j, k)] = c * rk + s * a[k];
            a[k] = -s * rk + c * a[k];
        }
        const double zj = st.z[j];
        st.z[j] = c * zj + s * y;
        y = -s * zj + c * y;
    }
}

__host__ __device__ __forceinline__ QrState qr_merge(QrState a, const QrState& b) {
#pragma unroll
    for (int i = 0; i < kLsmBasis; ++i) { // stream b's rows [R_b | z_b] into a
        double row[4];
#pragma unroll
        for (int j = 0; j < kLsmBasis; ++j) row[j] = j >= i ? b.r[ridx(i, j)] : 0.0;
        qr_add_row(a, row[0], row[1], row[2], row[3], b.z[i]);
    }
    a.rows += b.rows;
    return a;
}

__host__ __device__ __forceinline__ void laguerre4(double x, double* out) { // lsm.cpp:21-25
    const double e = exp(-x / 2.0);
    out[0] = e;
    out[1] = e * (1.0 - x);
    out[2] = e * (1.0 - 2.0 * x + 0.5 * x * x);
    out[3] = e * (1.0 - 3.0 * x + 1.5 * x * x - x * x * x / 6.0);
}

__host__ __device__ __forceinline__ double intrinsic_v(double s, double k, int type) {
    const double phi = type == 0 ? 1.0 : -1.0;
    const double v = phi * (s - k);
    return (v < 0.0) ? 0.0 : v;
}

struct LsmArgs {
    double spot, strike, disc, dt;
    double drift, vsd;     // FP64: per-step GBM constants (gbm.cpp:8-10 with dt)
    float mu2, sig2;       // FP32: log2-space step
    uint64_t seed, n_paths;
    int32_t m;             // exercise dates (monitoring points)
    uint8_t type;
    void* store;           // [date 1..M][path], float or double
    double* cf;            // [path]
    double* coeffs;        // [(M+1) x 4] device
    uint8_t* has;          // [M+1] device
    QrState* qr_partials;  // [blocks]
    Welford* partials;     // [tiles]
};

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
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    const uint32_t plo = (uint32_t)p, phh = (uint32_t)(p >> 32);
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

// cf = intrinsic at maturity (lsm.cpp:76-81).
template <bool FP64>
__global__ __launch_bounds__(256) void lsm_init_cf_kernel(LsmArgs a) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_paths) return;
    a.cf[p] = intrinsic_v(store_get<FP64>(a, a.m, p), a.strike, a.type);
}

__device__ __forceinline__ void qr_shfl_down(const QrState& s, QrState& o, int off) {
#pragma unroll
    for (int i = 0; i < 10; ++i) o.r[i] = __shfl_down(s.r[i], off, 64);
#pragma unroll
    for (int i = 0; i < 4; ++i) o.z[i] = __shfl_down(s.z[i], off, 64);
    o.rows = __shfl_down(s.rows, off, 64);
}

// Apply the exercise decision of date `date + 1` (if it has coefficients), discount, and
// fold this date's ITM rows. With `regress` false (the final pass, date 0) only the
// decision and the discount run. Grid: one lane per path, blocks of 256, every block
// writes its QR partial.
template <bool FP64>
__global__ __launch_bounds__(kLsmBlock) void lsm_date_kernel(LsmArgs a, int date, int regress) {
    __shared__ QrState lds[4];
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    QrState st = {};
    if (p < a.n_paths) {
        double cf = a.cf[p];
        const int later = date + 1;
        if (later <= a.m - 1 && a.has[later]) { // lsm.cpp:110-122 at the later date
            const double s1 = store_get<FP64>(a, later, p);
            const double ex = intrinsic_v(s1, a.strike, a.type);
            if (ex > 0.0) {
                double b[4];
                laguerre4(s1 / a.strike, b);
                const double* beta = a.coeffs + 4 * later;
                double cont = 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) cont += beta[j] * b[j];
                if (ex > cont) cf = ex;
            }
        }
        cf *= a.disc; // lsm.cpp:84-86 (and :125-127 after the last date)
        a.cf[p] = cf;
        if (regress) {
            const double s0 = store_get<FP64>(a, date, p);
            if (intrinsic_v(s0, a.strike, a.type) > 0.0) { // lsm.cpp:88-100
                double b[4];
                laguerre4(s0 / a.strike, b);
                qr_add_row(st, b[0], b[1], b[2], b[3], cf);
                st.rows = 1;
            }
        }
    }
    if (!regress) return; // uniform across the grid
    const int lane = threadIdx.x & 63;
#pragma unroll 1
    for (int off = 1; off < 64; off <<= 1) {
        QrState o;
        qr_shfl_down(st, o, off);
        if ((lane & (2 * off - 1)) == 0) st = qr_merge(st, o);
    }
    if (lane == 0) lds[threadIdx.x >> 6] = st;
    __syncthreads();
    if (threadIdx.x == 0) {
        QrState t = qr_merge(qr_merge(lds[0], lds[1]), qr_merge(lds[2], lds[3]));
        a.qr_partials[blockIdx.x] = t;
    }
}

// One block: merge the block partials along a fixed tree (1024 groups by the static
// partition, then pairwise), then back-substitute (linalg.cpp:58-69) when at least four
// ITM rows were folded (lsm.cpp:102-104).
__global__ __launch_bounds__(1024) void lsm_solve_kernel(LsmArgs a, int date, uint64_t n_blocks) {
    __shared__ QrState lds[256];
    const uint32_t g = threadIdx.x;
    // 256 groups (QrState is large; 256 x 116 B fits LDS comfortably)
    QrState acc = {};
    if (g < 256) {
        const uint64_t q = n_blocks / 256, rem = n_blocks % 256;
        const uint64_t b = g * q + (g < rem ? g : rem);
        const uint64_t e = b + q + (g < rem ? 1 : 0);
        for (uint64_t i = b; i < e; ++i) acc = qr_merge(acc, a.qr_partials[i]);
        lds[g] = acc;
    }
    __syncthreads();
    for (uint32_t s = 1; s < 256; s <<= 1) {
        if (g < 256 && (g & (2 * s - 1)) == 0) lds[g] = qr_merge(lds[g], lds[g + s]);
        __syncthreads();
    }
    if (g == 0) {
        const QrState t = lds[0];
        if (t.rows < (uint32_t)kLsmBasis) {
            a.has[date] = 0;
            return;
        }
        double beta[4] = {0, 0, 0, 0};
        for (int i = kLsmBasis - 1; i >= 0; --i) {
            double sum = t.z[i];
            for (int j = i + 1; j < kLsmBasis; ++j) sum -= t.r[ridx(i, j)] * beta[j];
            const double diag = t.r[ridx(i, i)];
            beta[i] = diag != 0.0 ? sum / diag : 0.0;
        }
        for (int j = 0; j < 4; ++j) a.coeffs[4 * date + j] = beta[j];
        a.has[date] = 1;
    }
}

// Welford over the final cash flows in the canonical tile order.
__global__ __launch_bounds__(kTileLanes) void lsm_welford_kernel(LsmArgs a) {
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

__global__ void lsm_clear_policy_kernel(uint8_t* has, double* coeffs, int m) {
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i <= m) {
        has[i] = 0;
        for (int j = 0; j < 4; ++j) coeffs[4 * i + j] = 0.0;
    }
}

namespace {

struct LsmRun {
    LsmArgs a;
    uint64_t blocks, tiles;
};

int lsm_setup(double spot, double strike, double rate, double carry, double vol, double time,
              uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision, ThreadCtx* ctx,
              LsmRun* run) {
    if (type > 1 || m < 1 || n == 0 || n > (1ull << 30)) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    const bool fp64 = precision == MCDX_PREC_FP64;
    LsmArgs a{};
    const double dt = time / (double)m;
    a.spot = spot;
    a.strike = strike;
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
    const size_t store_bytes = (size_t)m * n * (fp64 ? 8 : 4);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_store = take(store_bytes), o_cf = take(n * 8),
                 o_coeffs = take((size_t)(m + 1) * 4 * 8), o_has = take((size_t)m + 1),
                 o_qr = take(blocks * sizeof(QrState)), o_part = take((tiles + 1) * sizeof(Welford));
    int st = ensure_dev(ctx, off);
    if (st) return st;
    char* base = (char*)ctx->dev;
    a.store = base + o_store;
    a.cf = (double*)(base + o_cf);
    a.coeffs = (double*)(base + o_coeffs);
    a.has = (uint8_t*)(base + o_has);
    a.qr_partials = (QrState*)(base + o_qr);
    a.partials = (Welford*)(base + o_part);
    run->a = a;
    run->blocks = blocks;
    run->tiles = tiles;
    return MCDX_OK;
}

template <bool FP64>
hipError_t enqueue_paths(const LsmRun& r, hipStream_t s) {
    const unsigned g = (unsigned)r.blocks;
    hipLaunchKernelGGL(lsm_paths_kernel<FP64>, dim3(g), dim3(256), 0, s, r.a);
    hipLaunchKernelGGL(lsm_init_cf_kernel<FP64>, dim3(g), dim3(256), 0, s, r.a);
    hipLaunchKernelGGL(lsm_clear_policy_kernel, dim3((r.a.m + 1 + 255) / 256), dim3(256), 0, s,
                       r.a.has, r.a.coeffs, r.a.m);
    return hipGetLastError();
}

// Backward induction with regression (fit) or with a given policy already in coeffs/has.
template <bool FP64>
hipError_t enqueue_backward(const LsmRun& r, bool fit, hipStream_t s) {
    const unsigned g = (unsigned)r.blocks;
    for (int date = r.a.m - 1; date >= 1; --date) {
        hipLaunchKernelGGL(lsm_date_kernel<FP64>, dim3(g), dim3(kLsmBlock), 0, s, r.a, date, fit ? 1 : 0);
        if (fit)
            hipLaunchKernelGGL(lsm_solve_kernel, dim3(1), dim3(1024), 0, s, r.a, date, r.blocks);
    }
    hipLaunchKernelGGL(lsm_date_kernel<FP64>, dim3(g), dim3(kLsmBlock), 0, s, r.a, 0, 0);
    hipLaunchKernelGGL(lsm_welford_kernel, dim3((unsigned)r.tiles), dim3(kTileLanes), 0, s, r.a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_merge(r.a.partials, nullptr, r.tiles, 1, r.a.partials + r.tiles, s);
}

} // namespace

int lsm_price(double spot, double strike, double rate, double carry, double vol, double time,
              uint8_t type, int32_t m, uint64_t n, uint64_t seed, int precision, mcdx_result* out,
              double* coeffs_out, uint8_t* has_out, uint8_t* inception_out) {
    if (!out) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    LsmRun r;
    if ((st = lsm_setup(spot, strike, rate, carry, vol, time, type, m, n, seed, precision, ctx, &r)))
        return st;
    const size_t pol_bytes = (size_t)(m + 1) * 4 * 8 + (size_t)(m + 1);
    if ((st = ensure_pinned(ctx, sizeof(Welford) + pol_bytes + 64))) return st;
    const bool fp64 = precision == MCDX_PREC_FP64;
    hipError_t e = fp64 ? enqueue_paths<true>(r, ctx->stream) : enqueue_paths<false>(r, ctx->stream);
    if (e == hipSuccess)
        e = fp64 ? enqueue_backward<true>(r, true, ctx->stream) : enqueue_backward<false>(r, true, ctx->stream);
    char* h = (char*)ctx->pinned;
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, r.a.partials + r.tiles, sizeof(Welford), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h + 64, r.a.coeffs, (size_t)(m + 1) * 32, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h + 64 + (size_t)(m + 1) * 32, r.a.has, (size_t)(m + 1),
                           hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    const mcdx_result res = finish(*(const Welford*)h);
    const double* hc = (const double*)(h + 64);
    const uint8_t* hh = (const uint8_t*)(h + 64 + (size_t)(m + 1) * 32);
    if (coeffs_out)
        for (int i = 0; i < (m + 1) * 4; ++i) coeffs_out[i] = hc[i];
    if (has_out)
        for (int i = 0; i <= m; ++i) has_out[i] = hh[i];
    // inception decision (lsm.cpp:141-148)
    const double inc = intrinsic_v(spot, strike, type);
    if (inc > res.price) {
        if (inception_out) *inception_out = 1;
        *out = mcdx_result{inc, 0.0, res.path_count};
    } else {
        if (inception_out) *inception_out = 0;
        *out = res;
    }
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

} // namespace mcdx

using namespace mcdx;

extern "C" {

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

} // extern "C"
