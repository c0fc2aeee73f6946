// Fused GBM path kernels: Philox draw -> inverse normal -> exact GBM step -> O(1) payoff
// state -> FP64 Welford, one path per lane at a time, nothing stored per path.
//
// Reference behaviour restated here:
//   monte_carlo_terminal          src/pricers/monte_carlo.cpp:136-149 (European, digital)
//   monte_carlo_path_dependent    src/pricers/monte_carlo.cpp:155-177 (Asian, barrier,
//                                 lookback), draw index = step
//   barrier Brownian-bridge       src/pricers/monte_carlo.cpp:184-230 (stream-1 uniform)
//   arithmetic-Asian control var. src/pricers/monte_carlo.cpp:276-303
//   antithetic pair average       src/pricers/monte_carlo.cpp:73-80
//   payoffs                       include/mcd/payoffs/{european,digital,asian,barrier,
//                                 lookback}.hpp
//
// FP32 generation keeps the path in log2 space, x = log2(S/S0), advanced by one FMA per
// step, so barrier checks and running extrema compare x directly and only the products
// that need S inside the loop (arithmetic averages) pay an exp2 per step. Payoffs and all
// accumulation are FP64. FP64 mode repeats the reference's multiplicative update
// S <- S * exp(drift + vol*sqrt(dt)*z) with no contraction, step for step.
#include <math.h>

#include "runtime.hpp"

namespace mcdx {

enum Kind : int { kTerminal = 0, kAsian = 1, kBarrier = 2, kBarrierBB = 3, kLookback = 4 };

struct GbmArgs {
    double spot, strike, barrier, rebate, cash, discount;
    double drift, vsd;              // FP64: drift*dt and vol*sqrt(dt) (dt = T for terminal)
    double log_barrier, bb_denom;   // FP64 bridge: log(H), vol*vol*dt
    double kv;                      // Kemna-Vorst price for the control variate
    float mu2, sig2;                // FP32: drift*dt*log2(e), vol*sqrt(dt)*log2(e)
    float xb;                       // FP32: log2(H / S0)
    float bb_c;                     // FP32 bridge exponent scale: -2 ln2 / (vol^2 dt)
    uint64_t seed, n_paths, path0;
    int32_t m;
    uint8_t type, direction, knock, strike_style, average_style, digital_style, cv;
    Welford* partials;
    double* samples;
};

__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ double terminal_payoff(const GbmArgs& a, double st) {
    const double phi = a.type == 0 ? 1.0 : -1.0;
    if (a.digital_style == 0xFF) return smax(phi * (st - a.strike), 0.0); // european.hpp:19-22
    const bool itm = phi * (st - a.strike) > 0.0;                            // digital.hpp:16-23
    return !itm ? 0.0 : (a.digital_style == 0 ? a.cash : st);
}

// Barrier settlement: barrier.hpp:32-37 / monte_carlo.cpp:224-228.
__device__ __forceinline__ double barrier_payoff(const GbmArgs& a, double last, bool breached) {
    const double phi = a.type == 0 ? 1.0 : -1.0;
    const double vanilla = smax(phi * (last - a.strike), 0.0);
    const bool pays = a.knock == 1 ? !breached : breached;
    return pays ? vanilla : a.rebate;
}

// Asian settlement: asian.hpp:29-35 (arithmetic), :58-64 (geometric).
__device__ __forceinline__ double asian_payoff(const GbmArgs& a, double avg, double last) {
    const double phi = a.type == 0 ? 1.0 : -1.0;
    const double ref = a.strike_style == 0 ? (avg - a.strike) : (last - avg);
    return smax(phi * ref, 0.0);
}

// Lookback settlement: lookback.hpp:32-35 (fixed), :54-56 (floating).
__device__ __forceinline__ double lookback_payoff(const GbmArgs& a, double last, double mx,
                                                  double mn) {
    if (a.strike_style == 0)
        return a.type == 0 ? smax(mx - a.strike, 0.0) : smax(a.strike - mn, 0.0);
    return a.type == 0 ? (last - mn) : (mx - last);
}

// ---------------------------------------------------------------------------
// FP32 generation, FP64 payoff
// ---------------------------------------------------------------------------
struct StateF32 {
    float x;          // log2(S/S0)
    float xmax, xmin; // lookback extrema in log2 space
    float last_u;     // S/S0 at the last step (arithmetic Asian)
    double sum_u;     // sum of S/S0 over monitoring dates
    double sum_x;     // sum of log2(S/S0) (geometric average)
    bool breached;
};

template <int KIND>
__device__ __forceinline__ void observe_f32(const GbmArgs& a, StateF32& s, bool need_arith,
                                            bool need_geo) {
    if constexpr (KIND == kAsian) {
        if (need_arith) {
            const float u = fast_exp2(s.x);
            s.sum_u += (double)u;
            s.last_u = u;
        }
        if (need_geo) s.sum_x += (double)s.x;
    } else if constexpr (KIND == kBarrier) {
        const bool crossed = a.direction == 0 ? s.x >= a.xb : s.x <= a.xb;
        s.breached = s.breached || crossed;
    } else if constexpr (KIND == kLookback) {
        s.xmax = fmaxf(s.xmax, s.x);
        s.xmin = fminf(s.xmin, s.x);
    }
}

template <int KIND>
__device__ __forceinline__ double settle_f32(const GbmArgs& a, const StateF32& s) {
    if constexpr (KIND == kAsian) {
        const double last = a.spot * (double)(a.average_style == 0 || a.cv ? s.last_u : fast_exp2(s.x));
        const double arith = a.spot * (s.sum_u / (double)a.m);
        const double geo = a.spot * exp2(s.sum_x / (double)a.m);
        if (a.cv) { // monte_carlo.cpp:296-301
            const double phi = a.type == 0 ? 1.0 : -1.0;
            const double ap = a.discount * smax(phi * (arith - a.strike), 0.0);
            const double gp = a.discount * smax(phi * (geo - a.strike), 0.0);
            return ap - (gp - a.kv);
        }
        return a.discount * asian_payoff(a, a.average_style == 0 ? arith : geo, last);
    } else if constexpr (KIND == kBarrier || KIND == kBarrierBB) {
        return a.discount * barrier_payoff(a, a.spot * (double)fast_exp2(s.x), s.breached);
    } else {
        const double last = a.spot * (double)fast_exp2(s.x);
        const double mx = a.spot * (double)fast_exp2(s.xmax);
        const double mn = a.spot * (double)fast_exp2(s.xmin);
        return a.discount * lookback_payoff(a, last, mx, mn);
    }
}

template <int KIND, bool ANTI>
__device__ double path_f32(const GbmArgs& a, uint64_t path) {
#pragma clang fp contract(fast)
    const uint32_t plo = (uint32_t)path, phh = (uint32_t)(path >> 32);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    if constexpr (KIND == kTerminal) {
        const float z = inv_normal_f32(philox_w0(plo, phh, 0u, 0u, k0, k1));
        double r = a.discount * terminal_payoff(a, a.spot * (double)fast_exp2(fmaf(a.sig2, z, a.mu2)));
        if (ANTI) {
            const double r2 =
                a.discount * terminal_payoff(a, a.spot * (double)fast_exp2(fmaf(-a.sig2, z, a.mu2)));
            r = 0.5 * (r + r2);
        }
        return r;
    } else {
        const bool need_arith = a.average_style == 0 || a.cv;
        const bool need_geo = a.average_style == 1 || a.cv;
        StateF32 s = {0.f, 0.f, 0.f, 1.f, 0.0, 0.0, false};
        StateF32 t = s;
        // One monitoring step given its draw (draw index = step, monte_carlo.cpp:166-167).
        auto apply = [&](uint32_t draw, float z) {
            if constexpr (KIND == kBarrierBB) {
                // monte_carlo.cpp:200-223 in log2 space: the crossing probability
                // exp(-2 ln(H/S) ln(H/S') / (vol^2 dt)) = exp2(bb_c (xb - x)(xb - x')).
                const float xn = s.x + fmaf(a.sig2, z, a.mu2);
                const float xt = ANTI ? t.x + fmaf(-a.sig2, z, a.mu2) : 0.f;
                const bool cs = a.direction == 0 ? xn >= a.xb : xn <= a.xb;
                s.breached = s.breached || cs;
                if (ANTI) {
                    const bool ct = a.direction == 0 ? xt >= a.xb : xt <= a.xb;
                    t.breached = t.breached || ct;
                }
                if (!s.breached || (ANTI && !t.breached)) {
                    const uint32_t ub = philox_w0(plo, phh, draw, 1u, k0, k1);
                    const float u = fmaf((float)ub, 0x1p-32f, 0x1p-33f);
                    if (!s.breached) s.breached = u < fast_exp2(a.bb_c * (a.xb - s.x) * (a.xb - xn));
                    if (ANTI && !t.breached)
                        t.breached = u < fast_exp2(a.bb_c * (a.xb - t.x) * (a.xb - xt));
                }
                s.x = xn;
                if (ANTI) t.x = xt;
            } else {
                s.x = s.x + fmaf(a.sig2, z, a.mu2);
                observe_f32<KIND>(a, s, need_arith, need_geo);
                if (ANTI) {
                    t.x = t.x + fmaf(-a.sig2, z, a.mu2);
                    observe_f32<KIND>(a, t, need_arith, need_geo);
                }
            }
        };
        constexpr int B = 8;
        int step = 0;
        for (; step + B <= a.m; step += B) {
            float z[B];
            normal_block_f32<B>(plo, phh, (uint32_t)step, k0, k1, z);
#pragma unroll
            for (int j = 0; j < B; ++j) apply((uint32_t)(step + j), z[j]);
        }
        for (; step < a.m; ++step)
            apply((uint32_t)step, inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
        double r = settle_f32<KIND>(a, s);
        if (ANTI) r = 0.5 * (r + settle_f32<KIND>(a, t));
        return r;
    }
}

// ---------------------------------------------------------------------------
// FP64: the reference's arithmetic, operation for operation
// ---------------------------------------------------------------------------
struct StateF64 {
    double s, sum, log_sum, last, mx, mn;
    bool breached;
};

template <int KIND>
__device__ __forceinline__ void observe_f64(const GbmArgs& a, StateF64& st, bool need_arith,
                                            bool need_geo) {
    if constexpr (KIND == kAsian) {
        if (need_arith) st.sum += st.s;
        if (need_geo) st.log_sum += log(st.s);
        st.last = st.s;
    } else if constexpr (KIND == kBarrier) {
        const bool crossed = a.direction == 0 ? st.s >= a.barrier : st.s <= a.barrier;
        st.breached = st.breached || crossed;
        st.last = st.s;
    } else if constexpr (KIND == kLookback) {
        st.mx = smax(st.mx, st.s);
        st.mn = smin(st.mn, st.s);
        st.last = st.s;
    }
}

template <int KIND>
__device__ __forceinline__ double settle_f64(const GbmArgs& a, const StateF64& st) {
    if constexpr (KIND == kAsian) {
        const double count = (double)a.m;
        if (a.cv) {
            const double phi = a.type == 0 ? 1.0 : -1.0;
            const double ap = a.discount * smax(phi * (st.sum / count - a.strike), 0.0);
            const double gp = a.discount * smax(phi * (exp(st.log_sum / count) - a.strike), 0.0);
            return ap - (gp - a.kv);
        }
        const double avg = a.average_style == 0 ? st.sum / count : exp(st.log_sum / count);
        return a.discount * asian_payoff(a, avg, st.last);
    } else if constexpr (KIND == kBarrier) {
        return a.discount * barrier_payoff(a, st.last, st.breached);
    } else if constexpr (KIND == kBarrierBB) {
        return a.discount * barrier_payoff(a, st.s, st.breached);
    } else {
        return a.discount * lookback_payoff(a, st.last, st.mx, st.mn);
    }
}

// One bridge-corrected barrier step (monte_carlo.cpp:197-223); the stream-1 uniform for
// (path, step) is drawn at most once and shared by an antithetic pair, as both reference
// passes read the same counter.
__device__ __forceinline__ void bb_step_f64(const GbmArgs& a, StateF64& p, double z, double& u,
                                            bool& have_u, uint32_t plo, uint32_t phh,
                                            uint32_t draw, uint32_t k0, uint32_t k1) {
    const double sn = p.s * exp(a.drift + a.vsd * z);
    const bool crossed = a.direction == 0 ? sn >= a.barrier : sn <= a.barrier;
    p.breached = p.breached || crossed;
    if (!p.breached) {
        const double ls = log(p.s);
        const double lsn = log(sn);
        const double pc = exp(-2.0 * (a.log_barrier - ls) * (a.log_barrier - lsn) / a.bb_denom);
        if (!have_u) {
            u = uniform_open01_f64(philox_w0(plo, phh, draw, 1u, k0, k1));
            have_u = true;
        }
        p.breached = u < pc;
    }
    p.s = sn;
}

template <int KIND, bool ANTI>
__device__ double path_f64(const GbmArgs& a, uint64_t path) {
    const uint32_t plo = (uint32_t)path, phh = (uint32_t)(path >> 32);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    if constexpr (KIND == kTerminal) {
        const double z = inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, 0u, 0u, k0, k1)));
        double r = a.discount * terminal_payoff(a, a.spot * exp(a.drift + a.vsd * z));
        if (ANTI) {
            const double r2 = a.discount * terminal_payoff(a, a.spot * exp(a.drift + a.vsd * -z));
            r = 0.5 * (r + r2);
        }
        return r;
    } else {
        const bool need_arith = a.average_style == 0 || a.cv;
        const bool need_geo = a.average_style == 1 || a.cv;
        StateF64 s = {a.spot, 0.0, 0.0, a.spot, a.spot, a.spot, false};
        StateF64 t = s;
        for (int step = 0; step < a.m; ++step) {
            const uint32_t draw = (uint32_t)step;
            const double z =
                inv_normal_f64(uniform_open01_f64(philox_w0(plo, phh, draw, 0u, k0, k1)));
            if constexpr (KIND == kBarrierBB) {
                double u = 0.0;
                bool have_u = false;
                bb_step_f64(a, s, z, u, have_u, plo, phh, draw, k0, k1);
                if (ANTI) bb_step_f64(a, t, -z, u, have_u, plo, phh, draw, k0, k1);
            } else {
                s.s = s.s * exp(a.drift + a.vsd * z);
                observe_f64<KIND>(a, s, need_arith, need_geo);
                if (ANTI) {
                    t.s = t.s * exp(a.drift + a.vsd * -z);
                    observe_f64<KIND>(a, t, need_arith, need_geo);
                }
            }
        }
        double r = settle_f64<KIND>(a, s);
        if (ANTI) r = 0.5 * (r + settle_f64<KIND>(a, t));
        return r;
    }
}

template <int KIND, bool FP64, bool ANTI>
__device__ __forceinline__ double path_sample(const GbmArgs& a, uint64_t path) {
    if constexpr (FP64) return path_f64<KIND, ANTI>(a, path);
    else return path_f32<KIND, ANTI>(a, path);
}

// One tile of 1024 paths per block: lane folds its 4 paths (Welford, stats.hpp:23-29),
// then the block's pairwise tree writes the tile partial.
template <int KIND, bool FP64, bool ANTI>
__global__ __launch_bounds__(kTileLanes) void gbm_tile_kernel(GbmArgs a) {
    __shared__ Welford lds[4];
    const uint64_t tile = blockIdx.x;
    Welford w = {0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPathsPerLane; ++k) {
        const uint64_t path = tile * kTilePaths + (uint64_t)(k * kTileLanes) + threadIdx.x;
        if (path < a.n_paths) {
            const double x = path_sample<KIND, FP64, ANTI>(a, path);
            w.n += 1;
            const double d = x - w.mean;
            w.mean += d * (1.0 / (double)(k + 1));
            w.m2 += d * (x - w.mean);
        }
    }
    const Welford t = block_merge_256(w, lds);
    if (threadIdx.x == 0) a.partials[tile] = t;
}

template <int KIND, bool FP64, bool ANTI>
__global__ __launch_bounds__(256) void gbm_samples_kernel(GbmArgs a, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a.samples[i] = path_sample<KIND, FP64, ANTI>(a, a.path0 + i);
}

namespace {

int kind_of(const mcdx_gbm_contract& c) {
    switch (c.product) {
    case MCDX_EUROPEAN:
    case MCDX_DIGITAL: return kTerminal;
    case MCDX_ASIAN: return kAsian;
    case MCDX_BARRIER: return c.brownian_bridge ? kBarrierBB : kBarrier;
    case MCDX_LOOKBACK: return kLookback;
    default: return -1;
    }
}

int validate(const mcdx_gbm_contract* c, int precision) {
    if (!c) return MCDX_EINVAL;
    if (precision != MCDX_PREC_FP32 && precision != MCDX_PREC_FP64) return MCDX_EINVAL;
    if (kind_of(*c) < 0 || c->type > 1 || c->direction > 1 || c->knock > 1 ||
        c->strike_style > 1 || c->average_style > 1 || c->digital_style > 1)
        return MCDX_EINVAL;
    if (c->product != MCDX_EUROPEAN && c->product != MCDX_DIGITAL && c->monitoring_points < 1)
        return MCDX_EINVAL;
    if (c->path_count > (1ull << 40)) return MCDX_EINVAL;
    return MCDX_OK;
}

// Host precomputation mirrors the reference expressions exactly (gbm.cpp:8-10,
// monte_carlo.cpp:160-162,191-193, :276-283) so FP64 mode sees identical constants.
GbmArgs make_args(const mcdx_gbm_contract& c) {
    GbmArgs a{};
    const bool terminal = c.product == MCDX_EUROPEAN || c.product == MCDX_DIGITAL;
    const double dt = terminal ? c.time : c.time / (double)c.monitoring_points;
    a.spot = c.spot;
    a.strike = c.strike;
    a.barrier = c.barrier;
    a.rebate = c.rebate;
    a.cash = c.cash_amount;
    a.discount = exp(-c.rate * c.time);
    a.drift = (c.rate - c.carry_yield - 0.5 * c.vol * c.vol) * dt;
    a.vsd = c.vol * sqrt(dt);
    a.log_barrier = c.barrier > 0.0 ? log(c.barrier) : 0.0;
    a.bb_denom = c.vol * c.vol * dt;
    const double log2e = 1.44269504088896340735992468100189214;
    const double ln2 = 0.693147180559945309417232121458176568;
    a.mu2 = (float)(a.drift * log2e);
    a.sig2 = (float)(a.vsd * log2e);
    a.xb = c.barrier > 0.0 ? (float)log2(c.barrier / c.spot) : 0.f;
    a.bb_c = (float)(-2.0 * ln2 / a.bb_denom);
    a.seed = c.seed;
    a.n_paths = c.path_count;
    a.m = terminal ? 1 : c.monitoring_points;
    a.type = c.type;
    a.direction = c.direction;
    a.knock = c.knock;
    a.strike_style = c.strike_style;
    a.average_style = c.average_style;
    a.digital_style = c.product == MCDX_EUROPEAN ? 0xFF : c.digital_style;
    a.cv = (c.product == MCDX_ASIAN && c.average_style == 0 && c.control_variate &&
            c.strike_style == 0)
               ? 1
               : 0;
    if (a.cv) { // kemna_vorst, analytic.cpp:51-59 over BSM :30-38
        const double sigma_g = c.vol / 1.73205080756887729352744634150587237;
        const double b_g = 0.5 * (c.rate - c.carry_yield - c.vol * c.vol / 6.0);
        const double q_eff = c.rate - b_g;
        const double st = sqrt(c.time);
        const double d1 = (log(c.spot / c.strike) + (c.rate - q_eff + 0.5 * sigma_g * sigma_g) * c.time) /
                          (sigma_g * st);
        const double d2 = d1 - sigma_g * st;
        const double phi = c.type == 0 ? 1.0 : -1.0;
        const double ncdf1 = 0.5 * erfc(-(phi * d1) / 1.41421356237309504880168872420969808);
        const double ncdf2 = 0.5 * erfc(-(phi * d2) / 1.41421356237309504880168872420969808);
        a.kv = phi * c.spot * exp(-q_eff * c.time) * ncdf1 - phi * c.strike * exp(-c.rate * c.time) * ncdf2;
    }
    return a;
}

template <bool FP64, bool ANTI>
void launch_tiles_k(int kind, const GbmArgs& a, unsigned grid, hipStream_t s) {
    switch (kind) {
    case kTerminal: hipLaunchKernelGGL((gbm_tile_kernel<kTerminal, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    case kAsian: hipLaunchKernelGGL((gbm_tile_kernel<kAsian, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    case kBarrier: hipLaunchKernelGGL((gbm_tile_kernel<kBarrier, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    case kBarrierBB: hipLaunchKernelGGL((gbm_tile_kernel<kBarrierBB, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    default: hipLaunchKernelGGL((gbm_tile_kernel<kLookback, FP64, ANTI>), dim3(grid), dim3(kTileLanes), 0, s, a); break;
    }
}

template <bool FP64, bool ANTI>
void launch_samples_k(int kind, const GbmArgs& a, uint64_t n, hipStream_t s) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (kind) {
    case kTerminal: hipLaunchKernelGGL((gbm_samples_kernel<kTerminal, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kAsian: hipLaunchKernelGGL((gbm_samples_kernel<kAsian, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kBarrier: hipLaunchKernelGGL((gbm_samples_kernel<kBarrier, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    case kBarrierBB: hipLaunchKernelGGL((gbm_samples_kernel<kBarrierBB, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    default: hipLaunchKernelGGL((gbm_samples_kernel<kLookback, FP64, ANTI>), dim3(grid), dim3(256), 0, s, a, n); break;
    }
}

// Enqueue tiles + merge for one contract; scratch must hold tiles_for(N) + 1 Welfords.
hipError_t enqueue_contract(const mcdx_gbm_contract& c, int precision, Welford* partials,
                            Welford* result, hipStream_t s) {
    GbmArgs a = make_args(c);
    a.partials = partials;
    const uint64_t tiles = tiles_for(c.path_count);
    const int kind = kind_of(c);
    const bool fp64 = precision == MCDX_PREC_FP64;
    const bool anti = c.antithetic != 0;
    if (fp64) {
        if (anti) launch_tiles_k<true, true>(kind, a, (unsigned)tiles, s);
        else launch_tiles_k<true, false>(kind, a, (unsigned)tiles, s);
    } else {
        if (anti) launch_tiles_k<false, true>(kind, a, (unsigned)tiles, s);
        else launch_tiles_k<false, false>(kind, a, (unsigned)tiles, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_merge(partials, nullptr, tiles, 1, result, s);
}

} // namespace

// Shared with the batched book: price one contract synchronously on the thread's stream.
int price_gbm_sync(const mcdx_gbm_contract* c, int precision, mcdx_result* out) {
    int st = validate(c, precision);
    if (st) return st;
    if (!out) return MCDX_EINVAL;
    if (c->path_count == 0) {
        *out = mcdx_result{0.0, 0.0, 0};
        return MCDX_OK;
    }
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    const uint64_t tiles = tiles_for(c->path_count);
    if ((st = ensure_dev(ctx, (tiles + 1) * sizeof(Welford)))) return st;
    if ((st = ensure_pinned(ctx, sizeof(Welford)))) return st;
    Welford* partials = (Welford*)ctx->dev;
    hipError_t e = enqueue_contract(*c, precision, partials, partials + tiles, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->pinned, partials + tiles, sizeof(Welford), hipMemcpyDeviceToHost,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    *out = finish(*(const Welford*)ctx->pinned);
    return MCDX_OK;
}

} // namespace mcdx

using namespace mcdx;

namespace {
mcdx_gbm_contract base_contract(double spot, double strike, double rate, double carry,
                                double vol, double time, uint8_t type, uint64_t n,
                                uint64_t seed, const mcdx_options* o) {
    mcdx_gbm_contract c{};
    c.spot = spot;
    c.strike = strike;
    c.rate = rate;
    c.carry_yield = carry;
    c.vol = vol;
    c.time = time;
    c.cash_amount = 1.0;
    c.type = type;
    c.path_count = n;
    c.seed = seed;
    c.knock = 1;
    if (o) {
        c.antithetic = o->antithetic;
        c.control_variate = o->control_variate;
        c.brownian_bridge = o->brownian_bridge;
    }
    return c;
}
} // namespace

extern "C" {

int mcdx_price_gbm(const mcdx_gbm_contract* contract, int precision, mcdx_result* out) {
    return price_gbm_sync(contract, precision, out);
}

int mcdx_price_gbm_enqueue(const mcdx_gbm_contract* c, int precision, void* stream,
                           void* dev_welford) {
    int st = validate(c, precision);
    if (st) return st;
    if (!dev_welford) return MCDX_EINVAL;
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    hipStream_t s = pick_stream(ctx, stream);
    if (c->path_count == 0) {
        const Welford zero = {0, 0.0, 0.0};
        return status_of(hipMemcpyAsync(dev_welford, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    }
    const uint64_t tiles = tiles_for(c->path_count);
    // Scratch is reused by later calls on this thread; keep it stream-ordered on `s`.
    if (ctx->dev_bytes < tiles * sizeof(Welford)) {
        if (stream) {
            hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) return status_of(e);
        }
        if ((st = ensure_dev(ctx, tiles * sizeof(Welford)))) return st;
    }
    return status_of(enqueue_contract(*c, precision, (Welford*)ctx->dev, (Welford*)dev_welford, s));
}

int mcdx_gbm_path_samples(const mcdx_gbm_contract* c, int precision, uint64_t path0, uint64_t n,
                          double* out) {
    int st = validate(c, precision);
    if (st) return st;
    if (!out && n) return MCDX_EINVAL;
    if (n == 0) return MCDX_OK;
    if (n > (1ull << 28)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    if ((st = thread_ctx(&ctx))) return st;
    if ((st = ensure_dev(ctx, n * sizeof(double)))) return st;
    GbmArgs a = make_args(*c);
    a.samples = (double*)ctx->dev;
    a.path0 = path0;
    const int kind = kind_of(*c);
    const bool anti = c->antithetic != 0;
    if (precision == MCDX_PREC_FP64) {
        if (anti) launch_samples_k<true, true>(kind, a, n, ctx->stream);
        else launch_samples_k<true, false>(kind, a, n, ctx->stream);
    } else {
        if (anti) launch_samples_k<false, true>(kind, a, n, ctx->stream);
        else launch_samples_k<false, false>(kind, a, n, ctx->stream);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, ctx->dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return status_of(e);
}

int mcdx_monte_carlo_european(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint64_t path_count,
                              uint64_t seed, const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_EUROPEAN;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_digital(double spot, double strike, double rate, double carry_yield,
                             double vol, double time, uint8_t type, uint8_t style,
                             double cash_amount, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_DIGITAL;
    c.digital_style = style;
    c.cash_amount = cash_amount;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_asian(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, uint8_t strike_style,
                           uint8_t average_style, int32_t monitoring_points,
                           uint64_t path_count, uint64_t seed, const mcdx_options* options,
                           mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_ASIAN;
    c.strike_style = strike_style;
    c.average_style = average_style;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_barrier(double spot, double strike, double barrier, double rate,
                             double carry_yield, double vol, double time, uint8_t type,
                             uint8_t direction, uint8_t knock, double rebate,
                             int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                             const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_BARRIER;
    c.barrier = barrier;
    c.direction = direction;
    c.knock = knock;
    c.rebate = rebate;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

int mcdx_monte_carlo_lookback(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, uint8_t type, uint8_t style,
                              int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                              const mcdx_options* options, mcdx_result* out) {
    mcdx_gbm_contract c = base_contract(spot, strike, rate, carry_yield, vol, time, type,
                                        path_count, seed, options);
    c.product = MCDX_LOOKBACK;
    c.strike_style = style;
    c.monitoring_points = monitoring_points;
    return price_gbm_sync(&c, mcdx_get_precision(), out);
}

} // extern "C"
