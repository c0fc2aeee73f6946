// C-ABI wrapper over the reference `mcd` library, compiled together with the
// reference's own sources (oracle/Makefile) into oracle/_ref/libmcdref.so.
// TEST INFRASTRUCTURE ONLY: used to pin the C oracle and as the CPU baseline;
// the product library never links it. Reference sources are read in place and
// never copied into this repository.
#include "mcd/core/normal.hpp"
#include "mcd/core/rng.hpp"
#include "mcd/core/sobol.hpp"
#include "mcd/core/thread_pool.hpp"
#include "mcd/greeks/likelihood_ratio.hpp"
#include "mcd/greeks/pathwise.hpp"
#include "mcd/pricers/analytic.hpp"
#include "mcd/pricers/binomial.hpp"
#include "mcd/pricers/heston.hpp"
#include "mcd/pricers/lsm.hpp"
#include "mcd/pricers/monte_carlo.hpp"
#include "mcd/pricers/qmc.hpp"

#include <cstdint>

using namespace mcd;
using namespace mcd::pricers;

namespace {
struct RefOut {
    double price, standard_error;
    std::uint64_t path_count;
};
RefOut pack(const McResult& r) { return {r.price, r.standard_error, r.path_count}; }
McOptions opts(int anti, int cv, int bb, unsigned threads) {
    McOptions o;
    o.antithetic = anti != 0;
    o.control_variate = cv != 0;
    o.brownian_bridge = bb != 0;
    o.num_threads = threads;
    return o;
}
OptionType ot(int t) { return t == 0 ? OptionType::Call : OptionType::Put; }
} // namespace

extern "C" {

unsigned ref_chunk_count() { return ThreadPool::logical_chunk_count(); }

void ref_philox(const std::uint32_t* ctr, const std::uint32_t* key, std::uint32_t* out) {
    const PhiloxCounter c{ctr[0], ctr[1], ctr[2], ctr[3]};
    const PhiloxKey k{key[0], key[1]};
    const PhiloxCounter r = philox4x32_10(c, k);
    for (int i = 0; i < 4; ++i) out[i] = r[static_cast<std::size_t>(i)];
}
double ref_standard_normal_variate(std::uint64_t seed, std::uint64_t path, std::uint32_t draw) {
    return standard_normal_variate(seed, path, draw);
}
double ref_uniform_variate(std::uint64_t seed, std::uint64_t path, std::uint32_t draw) {
    return uniform_variate(seed, path, draw);
}
double ref_inverse_normal_cdf(double u) { return inverse_standard_normal_cdf(u); }
double ref_sobol_point(unsigned dim, std::uint64_t index) { return sobol_point(dim, index); }

void ref_mc_european(double s, double k, double r, double q, double v, double t, int type,
                     std::uint64_t n, std::uint64_t seed, int anti, unsigned threads, RefOut* out) {
    *out = pack(monte_carlo_european(s, k, r, q, v, t, ot(type), n, seed, opts(anti, 0, 0, threads)));
}
void ref_mc_digital(double s, double k, double r, double q, double v, double t, int type, int style,
                    double cash, std::uint64_t n, std::uint64_t seed, int anti, unsigned threads,
                    RefOut* out) {
    *out = pack(monte_carlo_digital(s, k, r, q, v, t, ot(type),
                                    style == 0 ? DigitalStyle::CashOrNothing : DigitalStyle::AssetOrNothing,
                                    cash, n, seed, opts(anti, 0, 0, threads)));
}
void ref_mc_asian(double s, double k, double r, double q, double v, double t, int type, int sstyle,
                  int astyle, int m, std::uint64_t n, std::uint64_t seed, int anti, int cv,
                  unsigned threads, RefOut* out) {
    *out = pack(monte_carlo_asian(s, k, r, q, v, t, ot(type),
                                  sstyle == 0 ? StrikeStyle::Fixed : StrikeStyle::Floating,
                                  astyle == 0 ? AverageStyle::Arithmetic : AverageStyle::Geometric, m,
                                  n, seed, opts(anti, cv, 0, threads)));
}
void ref_mc_barrier(double s, double k, double h, double r, double q, double v, double t, int type,
                    int dir, int knock, double rebate, int m, std::uint64_t n, std::uint64_t seed,
                    int anti, int bb, unsigned threads, RefOut* out) {
    *out = pack(monte_carlo_barrier(s, k, h, r, q, v, t, ot(type),
                                    dir == 0 ? BarrierDirection::Up : BarrierDirection::Down,
                                    knock == 0 ? BarrierKnock::In : BarrierKnock::Out, rebate, m, n,
                                    seed, opts(anti, 0, bb, threads)));
}
void ref_mc_lookback(double s, double k, double r, double q, double v, double t, int type, int style,
                     int m, std::uint64_t n, std::uint64_t seed, int anti, unsigned threads,
                     RefOut* out) {
    *out = pack(monte_carlo_lookback(s, k, r, q, v, t, ot(type),
                                     style == 0 ? StrikeStyle::Fixed : StrikeStyle::Floating, m, n,
                                     seed, opts(anti, 0, 0, threads)));
}
double ref_qmc_european(double s, double k, double r, double q, double v, double t, int type,
                        std::uint64_t n) {
    return qmc_sobol_european(s, k, r, q, v, t, ot(type), n).price;
}
double ref_qmc_asian(double s, double k, double r, double q, double v, double t, int type, int sstyle,
                     int m, std::uint64_t n) {
    return qmc_sobol_asian(s, k, r, q, v, t, ot(type),
                           sstyle == 0 ? StrikeStyle::Fixed : StrikeStyle::Floating, m, n)
        .price;
}
void ref_heston_qe(const double* p9, double k, int type, int steps, std::uint64_t n,
                   std::uint64_t seed, RefOut* out) {
    const models::HestonParams p{p9[0], p9[1], p9[2], p9[3], p9[4], p9[5], p9[6], p9[7], p9[8]};
    const HestonMcResult r = heston_qe_european(p, k, ot(type), steps, n, seed);
    *out = {r.price, r.standard_error, r.path_count};
}
double ref_heston_semi_analytic(const double* p9, double k, int type) {
    const models::HestonParams p{p9[0], p9[1], p9[2], p9[3], p9[4], p9[5], p9[6], p9[7], p9[8]};
    return heston_semi_analytic(p, k, ot(type));
}
void ref_lsm(double s, double k, double r, double q, double v, double t, int type, int m,
             std::uint64_t n, std::uint64_t seed, RefOut* out, double* coeffs, std::uint8_t* has,
             std::uint8_t* inception) {
    const LsmResult res = monte_carlo_lsm_american(s, k, r, q, v, t, ot(type), m, n, seed);
    *out = {res.price, res.standard_error, n};
    for (std::size_t d = 0; d < res.policy.coefficients_by_date.size(); ++d) {
        const auto& c = res.policy.coefficients_by_date[d];
        has[d] = c.empty() ? 0 : 1;
        for (std::size_t b = 0; b < c.size(); ++b) coeffs[d * 4 + b] = c[b];
    }
    *inception = res.policy.exercise_at_inception ? 1 : 0;
}
double ref_bsm(double s, double k, double r, double q, double v, double t, int type) {
    return black_scholes_merton(s, k, r, q, v, t, ot(type));
}
double ref_reiner_rubinstein(double s, double k, double h, double r, double q, double v, double t,
                             int type, int dir, int knock, double rebate) {
    return reiner_rubinstein(s, k, h, r, q, v, t, ot(type),
                             dir == 0 ? BarrierDirection::Up : BarrierDirection::Down,
                             knock == 0 ? BarrierKnock::In : BarrierKnock::Out, rebate);
}
double ref_kemna_vorst(double s, double k, double r, double q, double v, double t, int type) {
    return kemna_vorst(s, k, r, q, v, t, ot(type));
}
double ref_lookback_fixed(double s, double k, double r, double q, double v, double t, int type) {
    return lookback_fixed_strike(s, k, r, q, v, t, ot(type));
}
double ref_lookback_floating(double s, double r, double q, double v, double t, int type) {
    return lookback_floating_strike(s, r, q, v, t, ot(type));
}
double ref_crr_american(double s, double k, double r, double q, double v, double t, int steps,
                        int type) {
    return crr_binomial_american(s, k, r, q, v, t, steps, ot(type));
}
void ref_pathwise_european(double s, double k, double r, double q, double v, double t, int type,
                           std::uint64_t n, std::uint64_t seed, double* out6) {
    const auto g = mcd::greeks::pathwise_european(s, k, r, q, v, t, ot(type), n, seed);
    const double o[6] = {g.delta.value, g.delta.standard_error, g.vega.value,
                         g.vega.standard_error, g.rho.value, g.rho.standard_error};
    for (int i = 0; i < 6; ++i) out6[i] = o[i];
}
void ref_lr_european(double s, double k, double r, double q, double v, double t, int type,
                     std::uint64_t n, std::uint64_t seed, double* out10) {
    const auto g = mcd::greeks::likelihood_ratio_european(s, k, r, q, v, t, ot(type), n, seed);
    const double o[10] = {g.delta.value, g.delta.standard_error, g.gamma.value,
                          g.gamma.standard_error, g.vega.value, g.vega.standard_error,
                          g.rho.value, g.rho.standard_error, g.theta->value,
                          g.theta->standard_error};
    for (int i = 0; i < 10; ++i) out10[i] = o[i];
}
void ref_pathwise_asian(double s, double k, double r, double q, double v, double t, int type,
                        int sstyle, int astyle, int m, std::uint64_t n, std::uint64_t seed,
                        double* out6) {
    const auto g = mcd::greeks::pathwise_asian(
        s, k, r, q, v, t, ot(type), sstyle == 0 ? StrikeStyle::Fixed : StrikeStyle::Floating,
        astyle == 0 ? AverageStyle::Arithmetic : AverageStyle::Geometric, m, n, seed);
    const double o[6] = {g.delta.value, g.delta.standard_error, g.vega.value,
                         g.vega.standard_error, g.rho.value, g.rho.standard_error};
    for (int i = 0; i < 6; ++i) out6[i] = o[i];
}
void ref_lr_digital(double s, double k, double r, double q, double v, double t, int type,
                    int style, double cash, std::uint64_t n, std::uint64_t seed, double* out10) {
    const auto g = mcd::greeks::likelihood_ratio_digital(
        s, k, r, q, v, t, ot(type),
        style == 0 ? DigitalStyle::CashOrNothing : DigitalStyle::AssetOrNothing, cash, n, seed);
    const double o[10] = {g.delta.value, g.delta.standard_error, g.gamma.value,
                          g.gamma.standard_error, g.vega.value, g.vega.standard_error,
                          g.rho.value, g.rho.standard_error, g.theta->value,
                          g.theta->standard_error};
    for (int i = 0; i < 10; ++i) out10[i] = o[i];
}
// out8: delta, gamma, vega, rho (value, SE); returns 1 when theta is absent, as the
// reference documents for barriers.
int ref_lr_barrier(double s, double k, double h, double r, double q, double v, double t,
                   int type, int dir, int knock, double rebate, int m, std::uint64_t n,
                   std::uint64_t seed, double* out8) {
    const auto g = mcd::greeks::likelihood_ratio_barrier(
        s, k, h, r, q, v, t, ot(type), dir == 0 ? BarrierDirection::Up : BarrierDirection::Down,
        knock == 0 ? BarrierKnock::In : BarrierKnock::Out, rebate, m, n, seed);
    const double o[8] = {g.delta.value, g.delta.standard_error, g.gamma.value,
                         g.gamma.standard_error, g.vega.value, g.vega.standard_error,
                         g.rho.value, g.rho.standard_error};
    for (int i = 0; i < 8; ++i) out8[i] = o[i];
    return g.theta.has_value() ? 0 : 1;
}
// Policy as the oracle passes it: coeffs[(m+1) x 4], has[m+1] (1 = populated date).
double ref_reprice_frozen(double s, double k, double r, double q, double v, double t, int type,
                          int m, std::uint64_t n, std::uint64_t seed, const double* coeffs,
                          const std::uint8_t* has, std::uint8_t inception) {
    LsmPolicy policy;
    policy.coefficients_by_date.resize(static_cast<std::size_t>(m) + 1);
    for (int d = 0; d <= m; ++d)
        if (has[d])
            policy.coefficients_by_date[static_cast<std::size_t>(d)].assign(coeffs + 4 * d,
                                                                            coeffs + 4 * d + 4);
    policy.exercise_at_inception = inception != 0;
    return reprice_against_frozen_policy(s, k, r, q, v, t, ot(type), m, n, seed, policy);
}
double ref_black76(double f, double k, double r, double v, double t, int type) {
    return black76(f, k, r, v, t, ot(type));
}

} // extern "C"
