// The drop-in C++ API's Monte Carlo functions, each a thin call into the device engine's C
// ABI (include/mcdx.h). The reference functions are noexcept and never report errors; on a
// device error these return NaN prices with path_count 0 instead (mcdx_last_status() tells
// why).
#include <cmath>
#include <cstdint>
#include <limits>

#include "mcd/core/sobol.hpp"
#include "mcd/engine.hpp"
#include "mcd/greeks/finite_difference.hpp"
#include "mcd/greeks/likelihood_ratio.hpp"
#include "mcd/greeks/pathwise.hpp"
#include "mcd/pricers/heston.hpp"
#include "mcd/pricers/lsm.hpp"
#include "mcd/pricers/monte_carlo.hpp"
#include "mcd/pricers/qmc.hpp"
#include "mcdx.h"

namespace mcd {
namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
thread_local int t_last_status = MCDX_OK;

std::uint8_t u8(auto e) { return static_cast<std::uint8_t>(e); }

mcdx_options opts(const pricers::McOptions& o) {
    mcdx_options r{};
    r.antithetic = o.antithetic;
    r.control_variate = o.control_variate;
    r.brownian_bridge = o.brownian_bridge;
    r.num_threads = o.num_threads;
    return r;
}

pricers::McResult mc(int status, const mcdx_result& r) {
    t_last_status = status;
    if (status != MCDX_OK) return {kNaN, kNaN, 0};
    return {r.price, r.standard_error, r.path_count};
}

} // namespace

int last_engine_status() noexcept { return t_last_status; }

double sobol_point(unsigned dimension, std::uint64_t index) noexcept {
    std::uint32_t x = 0;
    t_last_status = mcdx_sobol_ints(dimension, index, 1, &x);
    if (t_last_status != MCDX_OK) return kNaN;
    return static_cast<double>(x) / 4294967296.0;
}

namespace pricers {

McResult monte_carlo_european(double s, double k, double r, double q, double v, double t,
                              OptionType type, std::uint64_t n, std::uint64_t seed) noexcept {
    return monte_carlo_european(s, k, r, q, v, t, type, n, seed, McOptions{});
}

McResult monte_carlo_european(double s, double k, double r, double q, double v, double t,
                              OptionType type, std::uint64_t n, std::uint64_t seed,
                              McOptions options) noexcept {
    mcdx_result res{};
    const mcdx_options o = opts(options);
    return mc(mcdx_monte_carlo_european(s, k, r, q, v, t, u8(type), n, seed, &o, &res), res);
}

McResult monte_carlo_digital(double s, double k, double r, double q, double v, double t,
                             OptionType type, DigitalStyle style, double cash, std::uint64_t n,
                             std::uint64_t seed, McOptions options) noexcept {
    mcdx_result res{};
    const mcdx_options o = opts(options);
    return mc(mcdx_monte_carlo_digital(s, k, r, q, v, t, u8(type), u8(style), cash, n, seed, &o, &res), res);
}

McResult monte_carlo_asian(double s, double k, double r, double q, double v, double t,
                           OptionType type, StrikeStyle ss, AverageStyle as, int m,
                           std::uint64_t n, std::uint64_t seed, McOptions options) noexcept {
    mcdx_result res{};
    const mcdx_options o = opts(options);
    return mc(mcdx_monte_carlo_asian(s, k, r, q, v, t, u8(type), u8(ss), u8(as), m, n, seed, &o, &res), res);
}

McResult monte_carlo_barrier(double s, double k, double h, double r, double q, double v, double t,
                             OptionType type, BarrierDirection d, BarrierKnock kn, double rebate,
                             int m, std::uint64_t n, std::uint64_t seed, McOptions options) noexcept {
    mcdx_result res{};
    const mcdx_options o = opts(options);
    return mc(mcdx_monte_carlo_barrier(s, k, h, r, q, v, t, u8(type), u8(d), u8(kn), rebate, m, n,
                                       seed, &o, &res),
              res);
}

McResult monte_carlo_lookback(double s, double k, double r, double q, double v, double t,
                              OptionType type, StrikeStyle style, int m, std::uint64_t n,
                              std::uint64_t seed, McOptions options) noexcept {
    mcdx_result res{};
    const mcdx_options o = opts(options);
    return mc(mcdx_monte_carlo_lookback(s, k, r, q, v, t, u8(type), u8(style), m, n, seed, &o, &res), res);
}

QmcResult qmc_sobol_european(double s, double k, double r, double q, double v, double t,
                             OptionType type, std::uint64_t n) noexcept {
    double p = kNaN;
    t_last_status = mcdx_qmc_sobol_european(s, k, r, q, v, t, u8(type), n, &p);
    return QmcResult{t_last_status == MCDX_OK ? p : kNaN};
}

QmcResult qmc_sobol_asian(double s, double k, double r, double q, double v, double t,
                          OptionType type, StrikeStyle ss, int m, std::uint64_t n) noexcept {
    double p = kNaN;
    t_last_status = mcdx_qmc_sobol_asian(s, k, r, q, v, t, u8(type), u8(ss), m, n, &p);
    return QmcResult{t_last_status == MCDX_OK ? p : kNaN};
}

HestonMcResult heston_qe_european(const models::HestonParams& p, double strike, OptionType type,
                                  int m, std::uint64_t n, std::uint64_t seed) noexcept {
    const mcdx_heston_params hp{p.spot, p.rate, p.carry_yield, p.v0, p.kappa, p.theta, p.xi,
                                p.rho, p.time};
    mcdx_result res{};
    t_last_status = mcdx_heston_qe_european(&hp, strike, u8(type), m, n, seed, &res);
    if (t_last_status != MCDX_OK) return {kNaN, kNaN, 0};
    return {res.price, res.standard_error, res.path_count};
}

LsmResult monte_carlo_lsm_american(double s, double k, double r, double q, double v, double t,
                                   OptionType type, int m, std::uint64_t n,
                                   std::uint64_t seed) noexcept {
    LsmResult out;
    if (m < 1) {
        out.price = out.standard_error = kNaN;
        return out;
    }
    std::vector<double> co((m + 1) * 4);
    std::vector<std::uint8_t> has(m + 1);
    std::uint8_t inc = 0;
    mcdx_result res{};
    t_last_status = mcdx_lsm_american(s, k, r, q, v, t, u8(type), m, n, seed, mcdx_get_precision(),
                                      &res, co.data(), has.data(), &inc);
    if (t_last_status != MCDX_OK) {
        out.price = out.standard_error = kNaN;
        return out;
    }
    out.price = res.price;
    out.standard_error = res.standard_error;
    out.policy.coefficients_by_date.resize(m + 1);
    for (int d = 0; d <= m; ++d)
        if (has[d]) out.policy.coefficients_by_date[d].assign(co.begin() + 4 * d, co.begin() + 4 * d + 4);
    out.policy.exercise_at_inception = inc != 0;
    return out;
}

double reprice_against_frozen_policy(double s, double k, double r, double q, double v, double t,
                                     OptionType type, int m, std::uint64_t n, std::uint64_t seed,
                                     const LsmPolicy& policy) noexcept {
    if (m < 1) return kNaN;
    std::vector<double> co((m + 1) * 4, 0.0);
    std::vector<std::uint8_t> has(m + 1, 0);
    for (int d = 0; d <= m && d < (int)policy.coefficients_by_date.size(); ++d) {
        const auto& c = policy.coefficients_by_date[d];
        if (c.size() == 4) {
            has[d] = 1;
            for (int j = 0; j < 4; ++j) co[4 * d + j] = c[j];
        }
    }
    mcdx_result res{};
    t_last_status = mcdx_lsm_reprice_frozen(s, k, r, q, v, t, u8(type), m, n, seed, co.data(),
                                            has.data(), policy.exercise_at_inception ? 1 : 0,
                                            mcdx_get_precision(), &res);
    return t_last_status == MCDX_OK ? res.price : kNaN;
}

} // namespace pricers

namespace greeks {

EuropeanGreeks finite_difference_european(double s, double k, double r, double q, double v,
                                          double t, OptionType type, std::uint64_t n,
                                          std::uint64_t seed, BumpSizes b) noexcept {
    mcdx_result p[9];
    t_last_status = mcdx_fd_european_prices(s, k, r, q, v, t, u8(type), n, seed, b.spot, b.vol,
                                            b.rate, b.time, p);
    if (t_last_status != MCDX_OK) return {kNaN, kNaN, kNaN, kNaN, kNaN};
    // finite_difference.cpp:56-63
    return EuropeanGreeks{
        .delta = (p[1].price - p[2].price) / (2.0 * b.spot),
        .gamma = (p[1].price - 2.0 * p[0].price + p[2].price) / (b.spot * b.spot),
        .vega = (p[3].price - p[4].price) / (2.0 * b.vol),
        .theta = -(p[7].price - p[8].price) / (2.0 * b.time),
        .rho = (p[5].price - p[6].price) / (2.0 * b.rate),
    };
}

namespace {
LrGreeks lr_from(int status, const mcdx_result* g, bool theta) {
    t_last_status = status;
    LrGreeks out;
    auto f = [&](int i) {
        return status == MCDX_OK ? LrGreeksResult{g[i].price, g[i].standard_error}
                                 : LrGreeksResult{kNaN, kNaN};
    };
    out.delta = f(0);
    out.gamma = f(1);
    out.vega = f(2);
    out.rho = f(3);
    if (theta) out.theta = f(4);
    return out;
}
PathwiseGreeks pw_from(int status, const mcdx_result* g) {
    t_last_status = status;
    auto f = [&](int i) {
        return status == MCDX_OK ? PathwiseGreeksResult{g[i].price, g[i].standard_error}
                                 : PathwiseGreeksResult{kNaN, kNaN};
    };
    return PathwiseGreeks{f(0), f(1), f(2)};
}
} // namespace

LrGreeks likelihood_ratio_european(double s, double k, double r, double q, double v, double t,
                                   OptionType type, std::uint64_t n, std::uint64_t seed) noexcept {
    mcdx_result g[5];
    return lr_from(mcdx_likelihood_ratio_european(s, k, r, q, v, t, u8(type), n, seed, g), g, true);
}

LrGreeks likelihood_ratio_digital(double s, double k, double r, double q, double v, double t,
                                  OptionType type, DigitalStyle style, double cash,
                                  std::uint64_t n, std::uint64_t seed) noexcept {
    mcdx_result g[5];
    return lr_from(mcdx_likelihood_ratio_digital(s, k, r, q, v, t, u8(type), u8(style), cash, n,
                                                 seed, g),
                   g, true);
}

LrGreeks likelihood_ratio_barrier(double s, double k, double h, double r, double q, double v,
                                  double t, OptionType type, BarrierDirection d, BarrierKnock kn,
                                  double rebate, int m, std::uint64_t n, std::uint64_t seed) noexcept {
    mcdx_result g[4];
    return lr_from(mcdx_likelihood_ratio_barrier(s, k, h, r, q, v, t, u8(type), u8(d), u8(kn),
                                                 rebate, m, n, seed, g),
                   g, false);
}

PathwiseGreeks pathwise_european(double s, double k, double r, double q, double v, double t,
                                 OptionType type, std::uint64_t n, std::uint64_t seed) noexcept {
    mcdx_result g[3];
    return pw_from(mcdx_pathwise_european(s, k, r, q, v, t, u8(type), n, seed, g), g);
}

PathwiseGreeks pathwise_asian(double s, double k, double r, double q, double v, double t,
                              OptionType type, StrikeStyle ss, AverageStyle as, int m,
                              std::uint64_t n, std::uint64_t seed) noexcept {
    mcdx_result g[3];
    return pw_from(mcdx_pathwise_asian(s, k, r, q, v, t, u8(type), u8(ss), u8(as), m, n, seed, g), g);
}

} // namespace greeks
} // namespace mcd
