// Heston quadratic-exponential (Andersen 2008) path kernel, many strikes per path pass.
//
// Reference behaviour restated here:
//   qe_variance_step     src/pricers/heston.cpp:91-116 (psi_c = 1.5; quadratic branch draws
//                        Z_v = Phi^-1(U) from the stream-1 uniform, exponential branch uses U)
//   heston_qe_european   src/pricers/heston.cpp:120-156 (K0..K4 central discretisation of
//                        log S; z_s from stream 0, U from stream 1, both at draw = step)
// The reference simulates every path again for each strike with the same draws, so one
// pass here yields exactly the per-strike samples of separate calls.
//
// FP32 mode: the draws, the variance and log-spot state and the branch arithmetic are FP32;
// the moment matching (m, s^2, psi and the psi <= psi_c switch) is FP64; payoffs and
// accumulation are FP64. FP64 mode repeats heston.cpp's arithmetic operation for operation.
#include <math.h>

#include <vector>

#include "runtime.hpp"

namespace mcdx {

constexpr int kMaxStrikes = 32;

struct HestonArgs {
    // FP64 constants, computed on the host exactly as heston.cpp:123-132 / :93-100 do.
    double v0, kappa, theta, xi, dt, e;      // e = exp(-kappa*dt)
    double drift, k0, k1, k2, k3, k4, discount, log_s0, spot;
    double sA, sB;                           // s^2 = v*sA + sB
    float f_drift_k0, f_k1, f_k2, f_k3, f_k4; // FP32 log-spot step constants
    uint64_t seed, n_paths, path0;
    int32_t steps, n_strikes;
    uint8_t type;
    const double* strikes; // device
    Welford* partials;     // [n_strikes][tiles]
    uint64_t tiles;
    double* samples;       // per-path discounted payoff at strikes[0] (parity hook)
};

__device__ __forceinline__ double smax_h(double a, double b) { return (a < b) ? b : a; }

// Terminal spot of one path: FP32 generation (draws, variance and log-spot state, branch
// arithmetic) with the QE moment matching (m, s^2, psi) and the psi <= psi_c switch in FP64.
// phh = path >> 32 (wave-uniform in the tile kernel, where it then stays in scalar registers).
__device__ double heston_path_f32(const HestonArgs& a, uint64_t path, uint32_t phh) {
    const uint32_t plo = (uint32_t)path;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    float v = (float)a.v0;
    float x = 0.f; // log(S/S0)
    for (int step = 0; step < a.steps; ++step) {
        const uint32_t draw = (uint32_t)step;
        const float zs = inv_normal_f32(philox_w0(plo, phh, draw, 0u, k0, k1));
        const uint32_t ub = philox_w0(plo, phh, draw, 1u, k0, k1);
        // moment matching in FP64 (heston.cpp:94-98); m >= theta (1 - e) > 0
        const double vd = (double)v;
        const double m = __builtin_fma(vd - a.theta, a.e, a.theta);
        const double s2 = __builtin_fma(vd, a.sA, a.sB);
        const double m2 = m * m;
        const double psi = s2 * fast_rcp_f64(m2);
        float vn;
        if (psi <= 1.5) { // heston.cpp:99-107
            const float inv_psi = (float)(m2 * fast_rcp_f64(s2));
            const float two = 2.0f * inv_psi;
            const float b2 = two - 1.0f + fast_sqrt(two) * fast_sqrt(two - 1.0f);
            const float aa = (float)m * fast_rcp(1.0f + b2);
            const float zv = inv_normal_f32(ub);
            const float b = fast_sqrt(b2);
            vn = aa * (b + zv) * (b + zv);
        } else { // heston.cpp:108-114
            const float p = (float)((psi - 1.0) * fast_rcp_f64(psi + 1.0));
            const float beta = (1.0f - p) * fast_rcp((float)m);
            const float u = __builtin_fmaf((float)ub, 0x1p-32f, 0x1p-33f);
            const float one_minus_u = __builtin_fmaf((float)(~ub), 0x1p-32f, 0x1p-33f);
            // log((1-p)/(1-u)) / beta = ln2 * (log2(1-p) - log2(1-u)) / beta
            vn = u <= p ? 0.f
                        : 0.69314718055994530942f * (fast_log2(1.0f - p) - fast_log2(one_minus_u)) *
                              fast_rcp(beta);
        }
        const float dv = fmaxf(__builtin_fmaf(a.f_k3, v, a.f_k4 * vn), 0.f);
        x = x + (a.f_drift_k0 + __builtin_fmaf(a.f_k1, v, a.f_k2 * vn) + fast_sqrt(dv) * zs);
        v = vn;
    }
    return a.spot * (double)fast_exp2(x * 1.44269504088896340736f);
}

// Terminal spot of one path, the reference's FP64 arithmetic (heston.cpp:136-148).
__device__ double heston_path_f64(const HestonArgs& a, uint64_t path, uint32_t phh) {
    const uint32_t plo = (uint32_t)path;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    double v = a.v0;
    double log_s = a.log_s0;
    for (int step = 0; step < a.steps; ++step) {
        const uint32_t draw = (uint32_t)step;
        const double zs = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, draw, 0u, k0, k1)));
        const double u = uniform_open01_f64(philox_w0(plo, phh, draw, 1u, k0, k1));
        const double e = a.e;
        const double m = a.theta + (v - a.theta) * e;
        const double s2 = v * a.xi * a.xi * e / a.kappa * (1.0 - e) +
                          a.theta * a.xi * a.xi / (2.0 * a.kappa) * (1.0 - e) * (1.0 - e);
        const double psi = s2 / (m * m);
        double vn;
        if (psi <= 1.5) {
            const double inv_psi = 1.0 / psi;
            const double b2 = 2.0 * inv_psi - 1.0 + sqrt(2.0 * inv_psi) * sqrt(2.0 * inv_psi - 1.0);
            const double aa = m / (1.0 + b2);
            const double zv = inv_normal_f64(u);
            const double b = sqrt(b2);
            vn = aa * (b + zv) * (b + zv);
        } else {
            const double p = (psi - 1.0) / (psi + 1.0);
            const double beta = (1.0 - p) / m;
            vn = u <= p ? 0.0 : log((1.0 - p) / (1.0 - u)) / beta;
        }
        const double dv = smax_h(a.k3 * v + a.k4 * vn, 0.0);
        log_s += a.drift + a.k0 + a.k1 * v + a.k2 * vn + sqrt(dv) * zs;
        v = vn;
    }
    return exp(log_s);
}

template <bool FP64>
__device__ __forceinline__ double heston_terminal(const HestonArgs& a, uint64_t path,
                                                 uint32_t phh) {
    if constexpr (FP64) return heston_path_f64(a, path, phh);
    else return heston_path_f32(a, path, phh);
}

// One tile of 1024 paths per block. Each lane keeps its 4 terminal spots, then every strike
// folds them into a Welford accumulator and merges along the block's pairwise tree.
template <bool FP64>
__global__ __launch_bounds__(kTileLanes) void heston_tile_kernel(HestonArgs a) {
    __shared__ Welford lds4[4];
    const uint64_t tile = blockIdx.x;
    double st[kPathsPerLane];
    bool valid[kPathsPerLane];
#pragma unroll
    for (int k = 0; k < kPathsPerLane; ++k) {
        // base is a multiple of 256 and threadIdx.x < 256: every lane's path >> 32 is base >> 32.
        const uint64_t base = tile * kTilePaths + (uint64_t)(k * kTileLanes);
        const uint64_t path = base + threadIdx.x;
        valid[k] = path < a.n_paths;
        st[k] = valid[k] ? heston_terminal<FP64>(a, path, (uint32_t)(base >> 32)) : 0.0;
    }
    const double phi = a.type == 0 ? 1.0 : -1.0;
    for (int s = 0; s < a.n_strikes; ++s) {
        const double K = a.strikes[s];
        Welford w = {0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < kPathsPerLane; ++k) {
            if (!valid[k]) continue;
            const double x = a.discount * smax_h(phi * (st[k] - K), 0.0);
            w.n += 1;
            const double d = x - w.mean;
            w.mean += d * (1.0 / (double)(k + 1));
            w.m2 += d * (x - w.mean);
        }
        const Welford t = block_merge_256_all(w, lds4);
        a.partials[(uint64_t)s * a.tiles + tile] = t; // every lane stores the same value
    }
}

template <bool FP64>
__global__ __launch_bounds__(256) void heston_samples_kernel(HestonArgs a, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double phi = a.type == 0 ? 1.0 : -1.0;
    const double st = heston_terminal<FP64>(a, a.path0 + i, (uint32_t)((a.path0 + i) >> 32));
    a.samples[i] = a.discount * smax_h(phi * (st - a.strikes[0]), 0.0);
}

namespace {

int make_args(const mcdx_heston_params* p, uint8_t type, int32_t steps, uint64_t n,
              uint64_t seed, HestonArgs* out) {
    if (!p || type > 1 || steps < 1 || n > (1ull << 40)) return MCDX_EINVAL;
    if (!(p->kappa > 0.0) || !(p->xi > 0.0) || !(p->theta >= 0.0) || !(p->v0 >= 0.0))
        return MCDX_EINVAL;
    HestonArgs a{};
    const double dt = p->time / (double)steps; // heston.cpp:123-132
    a.v0 = p->v0;
    a.kappa = p->kappa;
    a.theta = p->theta;
    a.xi = p->xi;
    a.dt = dt;
    a.e = exp(-p->kappa * dt);
    a.drift = (p->rate - p->carry_yield) * dt;
    a.k0 = -p->rho * p->kappa * p->theta * dt / p->xi;
    a.k1 = 0.5 * dt * (p->kappa * p->rho / p->xi - 0.5) - p->rho / p->xi;
    a.k2 = 0.5 * dt * (p->kappa * p->rho / p->xi - 0.5) + p->rho / p->xi;
    a.k3 = 0.5 * dt * (1.0 - p->rho * p->rho);
    a.k4 = a.k3;
    a.discount = exp(-p->rate * p->time);
    a.log_s0 = log(p->spot);
    a.spot = p->spot;
    a.sA = p->xi * p->xi * a.e / p->kappa * (1.0 - a.e);
    a.sB = p->theta * p->xi * p->xi / (2.0 * p->kappa) * (1.0 - a.e) * (1.0 - a.e);
    a.f_drift_k0 = (float)(a.drift + a.k0);
    a.f_k1 = (float)a.k1;
    a.f_k2 = (float)a.k2;
    a.f_k3 = (float)a.k3;
    a.f_k4 = (float)a.k4;
    a.seed = seed;
    a.n_paths = n;
    a.steps = steps;
    a.type = type;
    *out = a;
    return MCDX_OK;
}

} // namespace

int heston_price(const mcdx_heston_params* p, const double* strikes, int32_t n_strikes,
                 uint8_t type, int32_t steps, uint64_t n, uint64_t seed, int precision,
                 mcdx_result* out) {
    if (!strikes || !out || n_strikes < 1 || n_strikes > kMaxStrikes) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    HestonArgs a;
    int st = make_args(p, type, steps, n, seed, &a);
    if (st) return st;
    if (n == 0) {
        for (int s = 0; s < n_strikes; ++s) out[s] = mcdx_result{0.0, 0.0, 0};
        return MCDX_OK;
    }
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    const uint64_t tiles = tiles_for(n);
    const size_t part_bytes = (size_t)n_strikes * (tiles + 1) * sizeof(Welford);
    const size_t off_bytes = ((size_t)n_strikes + 1) * sizeof(uint64_t);
    const size_t k_bytes = (size_t)n_strikes * sizeof(double);
    if ((st = ensure_dev(ctx, part_bytes + off_bytes + k_bytes + 256))) return st;
    if ((st = ensure_pinned(ctx, (size_t)n_strikes * sizeof(Welford) + off_bytes + k_bytes))) return st;
    Welford* partials = (Welford*)ctx->dev;
    Welford* results = partials + (size_t)n_strikes * tiles;
    uint64_t* offs = (uint64_t*)((char*)ctx->dev + part_bytes);
    double* dk = (double*)((char*)offs + off_bytes);
    // stage offsets and strikes through pinned memory
    uint64_t* h_offs = (uint64_t*)ctx->pinned;
    double* h_k = (double*)((char*)ctx->pinned + off_bytes);
    hipError_t e = hipStreamSynchronize(ctx->stream); // pinned staging is reused
    for (int s = 0; s <= n_strikes; ++s) h_offs[s] = (uint64_t)s * tiles;
    for (int s = 0; s < n_strikes; ++s) h_k[s] = strikes[s];
    if (e == hipSuccess)
        e = hipMemcpyAsync(offs, h_offs, off_bytes + k_bytes, hipMemcpyHostToDevice, ctx->stream);
    a.strikes = dk;
    a.n_strikes = n_strikes;
    a.partials = partials;
    a.tiles = tiles;
    if (e == hipSuccess) {
        if (precision == MCDX_PREC_FP64)
            hipLaunchKernelGGL(heston_tile_kernel<true>, dim3((unsigned)tiles), dim3(kTileLanes), 0,
                               ctx->stream, a);
        else
            hipLaunchKernelGGL(heston_tile_kernel<false>, dim3((unsigned)tiles), dim3(kTileLanes), 0,
                               ctx->stream, a);
        e = launch_check();
    }
    if (e == hipSuccess)
        e = launch_final_merge(partials, offs, tiles, (uint32_t)n_strikes, results, ctx->stream);
    Welford* h_res = (Welford*)((char*)ctx->pinned + off_bytes + k_bytes);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream); // staging reads done
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_res, results, (size_t)n_strikes * sizeof(Welford),
                           hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    for (int s = 0; s < n_strikes; ++s) out[s] = finish(h_res[s]);
    return MCDX_OK;
}

} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_heston_qe_european(const mcdx_heston_params* params, double strike, uint8_t type,
                            int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                            mcdx_result* out) {
    return heston_price(params, &strike, 1, type, monitoring_points, path_count, seed,
                        mcdx_get_precision(), out);
}

int mcdx_heston_qe_strikes(const mcdx_heston_params* params, const double* strikes,
                           int32_t n_strikes, uint8_t type, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, int precision, mcdx_result* out) {
    return heston_price(params, strikes, n_strikes, type, monitoring_points, path_count, seed,
                        precision, out);
}

int mcdx_heston_path_samples(const mcdx_heston_params* params, double strike, uint8_t type,
                             int32_t monitoring_points, uint64_t seed, int precision,
                             uint64_t path0, uint64_t n, double* out) {
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (!out && n) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 28)) return MCDX_EINVAL;
    HestonArgs a;
    int st = make_args(params, type, monitoring_points, path0 + n, seed, &a);
    if (st) return st;
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    if ((st = ensure_dev(ctx, n * sizeof(double) + 64))) return st;
    double* dk = (double*)ctx->dev;
    a.samples = dk + 8;
    a.strikes = dk;
    a.path0 = path0;
    hipError_t e = hipMemcpyAsync(dk, &strike, sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (e == hipSuccess) {
        if (precision == MCDX_PREC_FP64)
            hipLaunchKernelGGL(heston_samples_kernel<true>, dim3(grid), dim3(256), 0, ctx->stream, a, n);
        else
            hipLaunchKernelGGL(heston_samples_kernel<false>, dim3(grid), dim3(256), 0, ctx->stream, a, n);
        e = launch_check();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, a.samples, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

} // extern "C"
