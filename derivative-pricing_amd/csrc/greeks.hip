// Monte Carlo Greeks on the device: pathwise, likelihood-ratio and common-random-number
// finite differences, each as one pass over the paths with several FP64 Welford
// accumulators per path (one per reported sensitivity).
//
// Reference behaviour restated here (FP64 arithmetic, reference expression order):
//   pathwise_european        src/greeks/pathwise.cpp:28-62
//   pathwise_asian           src/greeks/pathwise.cpp:64-147
//   likelihood_ratio_*       src/greeks/likelihood_ratio.cpp:34-99 (terminal: European and
//                            digital), :105-154 (barrier: first-step delta/gamma scores,
//                            summed vega/rho scores, no theta)
//   finite_difference        src/greeks/finite_difference.cpp:31-65: nine European pricings
//                            with common random numbers; fused here into one pass that
//                            evaluates the nine bumped terminal payoffs from each draw
// The reference accumulates each sensitivity with one sequential Welford; here each uses
// the canonical tile tree, so values agree to rounding (the tests bound it at 1e-12).
#include <math.h>

#include "runtime.hpp"

namespace mcdx {

enum GreekKind : int {
    kPwEuropean = 0, // delta, vega, rho
    kPwAsian = 1,    // delta, vega, rho
    kLrTerminal = 2, // delta, gamma, vega, rho, theta
    kLrBarrier = 3,  // delta, gamma, vega, rho
    kFdEuropean = 4  // prices at: base, S+, S-, vol+, vol-, r+, r-, T+, T-
};


struct GreekArgs {
    double spot, strike, barrier, rebate, rate, carry, vol, time, cash;
    double bump_spot, bump_vol, bump_rate, bump_time;
    uint64_t seed, n_paths;
    int32_t m;
    uint8_t type, direction, knock, strike_style, average_style, digital, digital_style;
    int32_t n_acc;
    Welford* partials; // [n_acc][tiles]
    uint64_t tiles;
};

__device__ __forceinline__ double gmax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ double gbm_spot(double s, double r, double q, double v, double t,
                                           double z) { // gbm.cpp:7-12
    const double drift = (r - q - 0.5 * v * v) * t;
    const double diffusion = v * sqrt(t) * z;
    return s * exp(drift + diffusion);
}

__device__ __forceinline__ double draw_z(const GreekArgs& a, uint64_t path, uint32_t step) {
    return inv_normal_f64(uniform_open01_f64(philox_bits(a.seed, path, step, 0)));
}

template <int KIND>
__device__ void greek_samples(const GreekArgs& a, uint64_t path, double* x) {
    const double phi = a.type == 0 ? 1.0 : -1.0;
    if constexpr (KIND == kPwEuropean) { // pathwise.cpp:38-58
        const double discount = exp(-a.rate * a.time);
        const double sqrt_t = sqrt(a.time);
        const double z = draw_z(a, path, 0);
        const double s_t = gbm_spot(a.spot, a.rate, a.carry, a.vol, a.time, z);
        const bool itm = phi * (s_t - a.strike) > 0.0;
        const double slope = itm ? phi : 0.0;
        const double h = discount * gmax(phi * (s_t - a.strike), 0.0);
        x[0] = discount * slope * (s_t / a.spot);
        x[1] = discount * slope * (s_t * (sqrt_t * z - a.vol * a.time));
        x[2] = discount * slope * (s_t * a.time) - a.time * h;
    } else if constexpr (KIND == kPwAsian) { // pathwise.cpp:76-143
        const double dt = a.time / (double)a.m;
        const double sqrt_dt = sqrt(dt);
        const double discount = exp(-a.rate * a.time);
        const double n = (double)a.m;
        double s = a.spot, cum_z = 0.0, sum_s = 0.0, log_sum = 0.0;
        double sum_s_dvol = 0.0, sum_dlns_dvol = 0.0, sum_s_drate = 0.0, sum_dlns_drate = 0.0;
        double last_s_dvol = 0.0, last_s_drate = 0.0;
        for (int step = 0; step < a.m; ++step) {
            const double z = draw_z(a, path, (uint32_t)step);
            cum_z += z;
            s = gbm_spot(s, a.rate, a.carry, a.vol, dt, z);
            const double i = (double)(step + 1);
            const double dlns_dvol = -a.vol * i * dt + sqrt_dt * cum_z;
            const double dlns_drate = i * dt;
            const double ds_dvol = s * dlns_dvol;
            const double ds_drate = s * dlns_drate;
            sum_s += s;
            log_sum += log(s);
            sum_s_dvol += ds_dvol;
            sum_dlns_dvol += dlns_dvol;
            sum_s_drate += ds_drate;
            sum_dlns_drate += dlns_drate;
            last_s_dvol = ds_dvol;
            last_s_drate = ds_drate;
        }
        const bool arith = a.average_style == 0;
        const double avg = arith ? sum_s / n : exp(log_sum / n);
        const double d_avg_dspot = avg / a.spot;
        const double d_avg_dvol = arith ? sum_s_dvol / n : avg * (sum_dlns_dvol / n);
        const double d_avg_drate = arith ? sum_s_drate / n : avg * (sum_dlns_drate / n);
        const double reference = a.strike_style == 0 ? (avg - a.strike) : (s - avg);
        const bool itm = phi * reference > 0.0;
        const double slope = itm ? phi : 0.0;
        const double h = discount * gmax(phi * reference, 0.0);
        double dr, vr, rr;
        if (a.strike_style == 0) {
            dr = slope * d_avg_dspot;
            vr = slope * d_avg_dvol;
            rr = slope * d_avg_drate;
        } else {
            dr = slope * (s / a.spot - d_avg_dspot);
            vr = slope * (last_s_dvol - d_avg_dvol);
            rr = slope * (last_s_drate - d_avg_drate);
        }
        x[0] = discount * dr;
        x[1] = discount * vr;
        x[2] = discount * rr - a.time * h;
    } else if constexpr (KIND == kLrTerminal) { // likelihood_ratio.cpp:39-71
        const double discount = exp(-a.rate * a.time);
        const double sqrt_t = sqrt(a.time);
        const double mu = a.rate - a.carry - 0.5 * a.vol * a.vol;
        const double z = draw_z(a, path, 0);
        const double s_t = gbm_spot(a.spot, a.rate, a.carry, a.vol, a.time, z);
        double pay;
        if (!a.digital) {
            pay = gmax(phi * (s_t - a.strike), 0.0);
        } else {
            const bool itm = phi * (s_t - a.strike) > 0.0;
            pay = !itm ? 0.0 : (a.digital_style == 0 ? a.cash : s_t);
        }
        const double h = discount * pay;
        x[0] = h * (z / (a.spot * a.vol * sqrt_t));
        x[1] = h * ((z * z - a.vol * sqrt_t * z - 1.0) / (a.spot * a.spot * a.vol * a.vol * a.time));
        x[2] = h * ((z * z - 1.0) / a.vol - sqrt_t * z);
        x[3] = h * (sqrt_t * z / a.vol - a.time);
        x[4] = h * (a.rate - ((z * z - 1.0) / (2.0 * a.time) + mu * z / (a.vol * sqrt_t)));
    } else if constexpr (KIND == kLrBarrier) { // likelihood_ratio.cpp:111-150
        const double dt = a.time / (double)a.m;
        const double sqrt_dt = sqrt(dt);
        const double discount = exp(-a.rate * a.time);
        double s = a.spot, z1 = 0.0, vega_sum = 0.0, rho_sum = 0.0, last = 0.0;
        bool breached = false;
        for (int step = 0; step < a.m; ++step) {
            const double z = draw_z(a, path, (uint32_t)step);
            if (step == 0) z1 = z;
            vega_sum += (z * z - 1.0) / a.vol - sqrt_dt * z;
            rho_sum += sqrt_dt * z / a.vol;
            s = gbm_spot(s, a.rate, a.carry, a.vol, dt, z);
            const bool crossed = a.direction == 0 ? s >= a.barrier : s <= a.barrier;
            breached = breached || crossed;
            last = s;
        }
        const double vanilla = gmax(phi * (last - a.strike), 0.0);
        const bool pays = a.knock == 1 ? !breached : breached;
        const double h = discount * (pays ? vanilla : a.rebate);
        x[0] = h * (z1 / (a.spot * a.vol * sqrt_dt));
        x[1] = h * ((z1 * z1 - a.vol * sqrt_dt * z1 - 1.0) / (a.spot * a.spot * a.vol * a.vol * dt));
        x[2] = h * vega_sum;
        x[3] = h * (rho_sum - a.time);
    } else { // kFdEuropean: the nine CRN pricings of finite_difference.cpp:44-54
        const double z = draw_z(a, path, 0);
        const double S[9] = {a.spot, a.spot + a.bump_spot, a.spot - a.bump_spot, a.spot, a.spot,
                             a.spot, a.spot, a.spot, a.spot};
        const double V[9] = {a.vol, a.vol, a.vol, a.vol + a.bump_vol, a.vol - a.bump_vol,
                             a.vol, a.vol, a.vol, a.vol};
        const double R[9] = {a.rate, a.rate, a.rate, a.rate, a.rate, a.rate + a.bump_rate,
                             a.rate - a.bump_rate, a.rate, a.rate};
        const double T[9] = {a.time, a.time, a.time, a.time, a.time, a.time, a.time,
                             a.time + a.bump_time, a.time - a.bump_time};
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double st = gbm_spot(S[i], R[i], a.carry, V[i], T[i], z);
            x[i] = exp(-R[i] * T[i]) * gmax(phi * (st - a.strike), 0.0);
        }
    }
}

template <int KIND, int NACC>
__global__ __launch_bounds__(kTileLanes) void greek_tile_kernel(GreekArgs a) {
    __shared__ Welford lds4[4];
    const uint64_t tile = blockIdx.x;
    Welford w[NACC];
#pragma unroll
    for (int i = 0; i < NACC; ++i) w[i] = Welford{0, 0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < kPathsPerLane; ++k) {
        const uint64_t path = tile * kTilePaths + (uint64_t)(k * kTileLanes) + threadIdx.x;
        if (path < a.n_paths) {
            double x[NACC];
            greek_samples<KIND>(a, path, x);
#pragma unroll
            for (int i = 0; i < NACC; ++i) {
                w[i].n += 1;
                const double d = x[i] - w[i].mean;
                w[i].mean += d / (double)w[i].n;
                w[i].m2 += d * (x[i] - w[i].mean);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NACC; ++i) {
        const Welford t = block_merge_256_all(w[i], lds4);
        a.partials[(uint64_t)i * a.tiles + tile] = t;
    }
}

namespace {

int run_greeks(int kind, GreekArgs a, mcdx_result* out) {
    static const int nacc_of[5] = {3, 3, 5, 4, 9};
    const int nacc = nacc_of[kind];
    if (a.n_paths == 0) {
        for (int i = 0; i < nacc; ++i) out[i] = mcdx_result{0.0, 0.0, 0};
        return MCDX_OK;
    }
    if (a.n_paths > (1ull << 40)) return MCDX_EINVAL;
    ThreadCtx* ctx;
    int st = thread_ctx(&ctx);
    if (st) return st;
    const uint64_t tiles = tiles_for(a.n_paths);
    const size_t part = (size_t)nacc * tiles * sizeof(Welford);
    const size_t offs_bytes = (size_t)(nacc + 1) * sizeof(uint64_t);
    if ((st = ensure_dev(ctx, part + nacc * sizeof(Welford) + offs_bytes + 512))) return st;
    if ((st = ensure_pinned(ctx, offs_bytes + nacc * sizeof(Welford) + 64))) return st;
    Welford* partials = (Welford*)ctx->dev;
    Welford* results = partials + (size_t)nacc * tiles;
    uint64_t* offs = (uint64_t*)(results + nacc);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    uint64_t* h_offs = (uint64_t*)ctx->pinned;
    for (int i = 0; i <= nacc; ++i) h_offs[i] = (uint64_t)i * tiles;
    if (e == hipSuccess) e = hipMemcpyAsync(offs, h_offs, offs_bytes, hipMemcpyHostToDevice, ctx->stream);
    a.partials = partials;
    a.tiles = tiles;
    a.n_acc = nacc;
    const dim3 g((unsigned)tiles), b(kTileLanes);
    if (e == hipSuccess) {
        switch (kind) {
        case kPwEuropean: hipLaunchKernelGGL((greek_tile_kernel<kPwEuropean, 3>), g, b, 0, ctx->stream, a); break;
        case kPwAsian: hipLaunchKernelGGL((greek_tile_kernel<kPwAsian, 3>), g, b, 0, ctx->stream, a); break;
        case kLrTerminal: hipLaunchKernelGGL((greek_tile_kernel<kLrTerminal, 5>), g, b, 0, ctx->stream, a); break;
        case kLrBarrier: hipLaunchKernelGGL((greek_tile_kernel<kLrBarrier, 4>), g, b, 0, ctx->stream, a); break;
        default: hipLaunchKernelGGL((greek_tile_kernel<kFdEuropean, 9>), g, b, 0, ctx->stream, a); break;
        }
        e = launch_check();
    }
    if (e == hipSuccess) e = launch_final_merge(partials, offs, tiles, (uint32_t)nacc, results, ctx->stream);
    Welford* h_res = (Welford*)((char*)ctx->pinned + 64 + offs_bytes);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_res, results, nacc * sizeof(Welford), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return status_of(e);
    for (int i = 0; i < nacc; ++i) out[i] = finish(h_res[i]);
    return MCDX_OK;
}

GreekArgs base_args(double spot, double strike, double rate, double carry, double vol,
                    double time, uint8_t type, uint64_t n, uint64_t seed) {
    GreekArgs a{};
    a.spot = spot;
    a.strike = strike;
    a.rate = rate;
    a.carry = carry;
    a.vol = vol;
    a.time = time;
    a.type = type;
    a.n_paths = n;
    a.seed = seed;
    a.cash = 1.0;
    a.m = 1;
    return a;
}

} // namespace
} // namespace mcdx

using namespace mcdx;

extern "C" {

int mcdx_pathwise_european(double spot, double strike, double rate, double carry_yield,
                           double vol, double time, uint8_t type, uint64_t path_count,
                           uint64_t seed, mcdx_result* out3) {
    if (!out3 || type > 1) return MCDX_EINVAL;
    return run_greeks(kPwEuropean, base_args(spot, strike, rate, carry_yield, vol, time, type,
                                             path_count, seed), out3);
}

int mcdx_pathwise_asian(double spot, double strike, double rate, double carry_yield, double vol,
                        double time, uint8_t type, uint8_t strike_style, uint8_t average_style,
                        int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                        mcdx_result* out3) {
    if (!out3 || type > 1 || strike_style > 1 || average_style > 1 || monitoring_points < 1)
        return MCDX_EINVAL;
    GreekArgs a = base_args(spot, strike, rate, carry_yield, vol, time, type, path_count, seed);
    a.strike_style = strike_style;
    a.average_style = average_style;
    a.m = monitoring_points;
    return run_greeks(kPwAsian, a, out3);
}

int mcdx_likelihood_ratio_european(double spot, double strike, double rate, double carry_yield,
                                   double vol, double time, uint8_t type, uint64_t path_count,
                                   uint64_t seed, mcdx_result* out5) {
    if (!out5 || type > 1) return MCDX_EINVAL;
    return run_greeks(kLrTerminal, base_args(spot, strike, rate, carry_yield, vol, time, type,
                                             path_count, seed), out5);
}

int mcdx_likelihood_ratio_digital(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, uint8_t type, uint8_t style,
                                  double cash_amount, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out5) {
    if (!out5 || type > 1 || style > 1) return MCDX_EINVAL;
    GreekArgs a = base_args(spot, strike, rate, carry_yield, vol, time, type, path_count, seed);
    a.digital = 1;
    a.digital_style = style;
    a.cash = cash_amount;
    return run_greeks(kLrTerminal, a, out5);
}

int mcdx_likelihood_ratio_barrier(double spot, double strike, double barrier, double rate,
                                  double carry_yield, double vol, double time, uint8_t type,
                                  uint8_t direction, uint8_t knock, double rebate,
                                  int32_t monitoring_points, uint64_t path_count, uint64_t seed,
                                  mcdx_result* out4) {
    if (!out4 || type > 1 || direction > 1 || knock > 1 || monitoring_points < 1) return MCDX_EINVAL;
    GreekArgs a = base_args(spot, strike, rate, carry_yield, vol, time, type, path_count, seed);
    a.barrier = barrier;
    a.direction = direction;
    a.knock = knock;
    a.rebate = rebate;
    a.m = monitoring_points;
    return run_greeks(kLrBarrier, a, out4);
}

int mcdx_fd_european_prices(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, uint8_t type, uint64_t path_count,
                            uint64_t seed, double bump_spot, double bump_vol, double bump_rate,
                            double bump_time, mcdx_result* out9) {
    if (!out9 || type > 1) return MCDX_EINVAL;
    GreekArgs a = base_args(spot, strike, rate, carry_yield, vol, time, type, path_count, seed);
    a.bump_spot = bump_spot;
    a.bump_vol = bump_vol;
    a.bump_rate = bump_rate;
    a.bump_time = bump_time;
    return run_greeks(kFdEuropean, a, out9);
}

} // extern "C"
