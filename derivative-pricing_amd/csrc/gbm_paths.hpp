// GBM path simulation shared by the per-contract kernels (gbm.hip) and the batched book
// (book.hip): contract arguments, the FP32-generation and FP64 per-path samplers, and the
// host-side contract validation / constant precomputation.
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
#pragma once

#include <math.h>

#include <type_traits>

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
    float xmax, xmin; // extrema in log2 space: lookback (seeded at S0), discrete barrier
                      // (monitoring dates only)
    float last_u;     // S/S0 at the last step (arithmetic Asian)
    double sum_u;     // sum of S/S0 over monitoring dates (flushed blocks)
    double sum_x;     // sum of log2(S/S0) (geometric average)
    bool breached;
    float blk_u;      // FP32 sum of S/S0 over the current block of up to 8 dates
};

// NEED_ARITH / NEED_GEO: which Asian sums the contract needs, as template arguments so the
// step loop carries no per-step test (a runtime flag here is if-converted into FP64 work on
// every step).
template <int KIND, bool NEED_ARITH, bool NEED_GEO>
__device__ __forceinline__ void observe_f32(const GbmArgs& a, StateF32& s) {
    if constexpr (KIND == kAsian) {
        if constexpr (NEED_ARITH) { // the block sum joins sum_u in flush_f32
            const float u = fast_exp2(s.x);
            s.blk_u += u;
            s.last_u = u;
        }
        if constexpr (NEED_GEO) s.sum_x += (double)s.x;
    } else if constexpr (KIND == kBarrier || KIND == kLookback) {
        // A discrete barrier is breached iff some monitoring date's x crosses xb, i.e. iff
        // the largest (up) or smallest (down) x over the dates does: the same comparisons on
        // the same values, taken once in settle_f32 instead of on every step.
        s.xmax = fmaxf(s.xmax, s.x);
        s.xmin = fminf(s.xmin, s.x);
    }
}

// Adds the block's FP32 sum of S/S0 (at most 8 dates: relative error a few 1e-7, below the
// FP32 generation's own) to the FP64 path sum: one conversion and FP64 add per block of dates
// instead of per date. Called after every block of 8 dates and after the remaining dates.
template <int KIND, bool NEED_ARITH>
__device__ __forceinline__ void flush_f32(StateF32& s) {
    if constexpr (KIND == kAsian && NEED_ARITH) {
        s.sum_u += (double)s.blk_u;
        s.blk_u = 0.f;
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
    } else if constexpr (KIND == kBarrier) {
        const bool breached = a.direction == 0 ? s.xmax >= a.xb : s.xmin <= a.xb;
        return a.discount * barrier_payoff(a, a.spot * (double)fast_exp2(s.x), breached);
    } else if constexpr (KIND == kBarrierBB) {
        return a.discount * barrier_payoff(a, a.spot * (double)fast_exp2(s.x), s.breached);
    } else {
        const double last = a.spot * (double)fast_exp2(s.x);
        const double mx = a.spot * (double)fast_exp2(s.xmax);
        const double mn = a.spot * (double)fast_exp2(s.xmin);
        return a.discount * lookback_payoff(a, last, mx, mn);
    }
}

// `phh` = path >> 32, passed separately so that a caller whose lanes share it (a block of 256
// aligned paths) can hand it over in a scalar register: the first Philox round's left word and
// the second round's M0 product then stay on the scalar unit.
// Which running sums an Asian path keeps: decided per path from the contract (kSumsAny) or
// fixed at compile time (arithmetic, geometric, both), which leaves one step loop in the
// kernel and fewer registers.
enum AsianSums : int { kSumsAny = 0, kSumsArith = 1, kSumsGeo = 2, kSumsBoth = 3 };

template <int KIND, bool ANTI, int SUMS = kSumsAny>
__device__ double path_f32(const GbmArgs& a, uint64_t path, uint32_t phh) {
#pragma clang fp contract(fast)
    const uint32_t plo = (uint32_t)path;
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
        if constexpr (KIND == kBarrier) { // the barrier is watched on monitoring dates only
            s.xmax = -__builtin_inff();
            s.xmin = __builtin_inff();
        }
        StateF32 t = s;
        // The step loop, instantiated per combination of the Asian sums it needs.
        auto steps = [&](auto arith_tag, auto geo_tag) {
        constexpr bool NA = decltype(arith_tag)::value, NG = decltype(geo_tag)::value;
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
                // u >= 2^-33, so a crossing probability below 2^-34 can never exceed it: the
                // uniform is drawn only when some alive path has exponent >= -34 (spots near
                // the barrier); the decisions are exactly those of drawing it every step.
                const float es = a.bb_c * (a.xb - s.x) * (a.xb - xn);
                const float et = ANTI ? a.bb_c * (a.xb - t.x) * (a.xb - xt) : -1e30f;
                const bool need_s = !s.breached && es >= -34.f;
                const bool need_t = ANTI && !t.breached && et >= -34.f;
                if (need_s || need_t) {
                    const uint32_t ub = philox_w0(plo, phh, draw, 1u, k0, k1);
                    const float u = fmaf((float)ub, 0x1p-32f, 0x1p-33f);
                    if (need_s) s.breached = u < fast_exp2(es);
                    if (need_t) t.breached = u < fast_exp2(et);
                }
                s.x = xn;
                if (ANTI) t.x = xt;
            } else {
                s.x = s.x + fmaf(a.sig2, z, a.mu2);
                observe_f32<KIND, NA, NG>(a, s);
                if (ANTI) {
                    t.x = t.x + fmaf(-a.sig2, z, a.mu2);
                    observe_f32<KIND, NA, NG>(a, t);
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
            flush_f32<KIND, NA>(s);
            if (ANTI) flush_f32<KIND, NA>(t);
        }
        for (; step < a.m; ++step)
            apply((uint32_t)step, inv_normal_f32(philox_w0(plo, phh, (uint32_t)step, 0u, k0, k1)));
        flush_f32<KIND, NA>(s);
        if (ANTI) flush_f32<KIND, NA>(t);
        };
        using T = std::true_type;
        using F = std::false_type;
        if constexpr (KIND == kAsian && SUMS == kSumsArith) {
            steps(T{}, F{});
        } else if constexpr (KIND == kAsian && SUMS == kSumsGeo) {
            steps(F{}, T{});
        } else if constexpr (KIND == kAsian && SUMS == kSumsBoth) {
            steps(T{}, T{});
        } else if constexpr (KIND == kAsian) {
            if (need_arith && need_geo) steps(T{}, T{});
            else if (need_arith) steps(T{}, F{});
            else steps(F{}, T{});
        } else {
            steps(F{}, F{});
        }
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
        const double ex = -2.0 * (a.log_barrier - ls) * (a.log_barrier - lsn) / a.bb_denom;
        // u >= 2^-33 > exp(-23.5): below that exponent u < pc cannot hold, so the uniform is
        // not needed (the decision is the one the reference makes after drawing it).
        if (ex >= -23.5) {
            const double pc = exp(ex);
            if (!have_u) {
                u = uniform_open01_f64(philox_w0(plo, phh, draw, 1u, k0, k1));
                have_u = true;
            }
            p.breached = u < pc;
        }
    }
    p.s = sn;
}

template <int KIND, bool ANTI>
__device__ double path_f64(const GbmArgs& a, uint64_t path, uint32_t phh) {
    const uint32_t plo = (uint32_t)path;
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

// One path's sample; phh must equal path >> 32.
template <int KIND, bool FP64, bool ANTI, int SUMS = kSumsAny>
__device__ __forceinline__ double path_sample(const GbmArgs& a, uint64_t path, uint32_t phh) {
    if constexpr (FP64) return path_f64<KIND, ANTI>(a, path, phh);
    else return path_f32<KIND, ANTI, SUMS>(a, path, phh);
}

// The Asian sums a contract needs (the same choice path_f32 makes per path).
__host__ __device__ inline int asian_sums(const GbmArgs& a) {
    const bool arith = a.average_style == 0 || a.cv;
    const bool geo = a.average_style == 1 || a.cv;
    return arith && geo ? kSumsBoth : (arith ? kSumsArith : kSumsGeo);
}


inline int kind_of(const mcdx_gbm_contract& c) {
    switch (c.product) {
    case MCDX_EUROPEAN:
    case MCDX_DIGITAL: return kTerminal;
    case MCDX_ASIAN: return kAsian;
    case MCDX_BARRIER: return c.brownian_bridge ? kBarrierBB : kBarrier;
    case MCDX_LOOKBACK: return kLookback;
    default: return -1;
    }
}

inline int validate(const mcdx_gbm_contract* c, int precision) {
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
inline GbmArgs make_args(const mcdx_gbm_contract& c) {
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


} // namespace mcdx
