// C++ tests of the drop-in API (include/mcd/**, libmcd.so), in the spirit of the reference's
// tests/*.cpp: closed forms against published or textbook values, Monte Carlo results
// against their closed forms within 3-4 standard errors, and the engine-error convention.
//
// Usage: api_test host   -- host-side functions only; Monte Carlo calls must report a
//                           device error as {NaN, NaN, 0} with a nonzero engine status
//        api_test device -- everything, on a GPU
// Exit status 0 when every check passes; each failure is printed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "mcd/core/sobol.hpp"
#include "mcd/core/types.hpp"
#include "mcd/engine.hpp"
#include "mcd/greeks/finite_difference.hpp"
#include "mcd/greeks/likelihood_ratio.hpp"
#include "mcd/greeks/pathwise.hpp"
#include "mcd/models/heston.hpp"
#include "mcd/pricers/analytic.hpp"
#include "mcd/pricers/binomial.hpp"
#include "mcd/pricers/heston.hpp"
#include "mcd/pricers/lsm.hpp"
#include "mcd/pricers/monte_carlo.hpp"
#include "mcd/pricers/qmc.hpp"

namespace {

int g_failures = 0;
int g_checks = 0;

void check(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        ++g_failures;
        std::printf("FAIL: %s\n", what.c_str());
    }
}

void near(double got, double want, double tol, const std::string& what) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: got %.12g want %.12g tol %.3g", what.c_str(), got, want, tol);
    check(std::fabs(got - want) <= tol, buf);
}

using mcd::OptionType;
namespace P = mcd::pricers;
namespace G = mcd::greeks;

void host_tests() {
    // Black-Scholes C1 value and put-call parity.
    const double c = P::black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call);
    const double p = P::black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Put);
    near(c, 10.450583572185565, 1e-12, "BSM call");
    near(c - p, 100.0 - 100.0 * std::exp(-0.05), 1e-12, "put-call parity");
    // Forward and Black-76 consistency: Black-76 on the forward equals BSM.
    const double f = P::forward_price(100, 0.05, 0.02, 1.0);
    near(f, 100.0 * std::exp(0.03), 1e-12, "forward price");
    near(P::black76(f, 100, 0.05, 0.2, 1.0, OptionType::Call),
         P::black_scholes_merton(100, 100, 0.05, 0.02, 0.2, 1.0, OptionType::Call), 1e-10, "black76");
    // Barrier in + out = vanilla.
    for (auto dir : {mcd::BarrierDirection::Up, mcd::BarrierDirection::Down}) {
        const double h = dir == mcd::BarrierDirection::Up ? 120.0 : 85.0;
        const double in = P::reiner_rubinstein(100, 100, h, 0.05, 0.0, 0.2, 1.0, OptionType::Call, dir,
                                               mcd::BarrierKnock::In);
        const double out = P::reiner_rubinstein(100, 100, h, 0.05, 0.0, 0.2, 1.0, OptionType::Call, dir,
                                                mcd::BarrierKnock::Out);
        near(in + out, c, 1e-9, "barrier in+out parity");
    }
    // Digital: cash + asset decomposition of the vanilla call.
    const double asset = P::digital(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                    mcd::DigitalStyle::AssetOrNothing);
    const double cash = P::digital(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                   mcd::DigitalStyle::CashOrNothing, 100.0);
    near(asset - cash, c, 1e-10, "asset - K x cash digital = call");
    // Kemna-Vorst below the arithmetic Asian's upper bound (the European).
    const double kv = P::kemna_vorst(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call);
    check(kv > 0.0 && kv < c, "Kemna-Vorst between 0 and the European");
    // CRR converges to BSM; American call on a non-dividend stock equals the European.
    near(P::crr_binomial(100, 100, 0.05, 0.0, 0.2, 1.0, 4000, OptionType::Call).price, c, 2e-3, "CRR -> BSM");
    near(P::crr_binomial_american(100, 100, 0.05, 0.0, 0.2, 1.0, 2000, OptionType::Call),
         P::crr_binomial(100, 100, 0.05, 0.0, 0.2, 1.0, 2000, OptionType::Call).price, 1e-9,
         "American call = European (q = 0)");
    // Heston semi-analytic: the published value the reference's CLI test uses.
    const mcd::models::HestonParams hp{100, 0.01, 0.02, 0.04, 4.0, 0.25, 1.0, -0.5, 1.0};
    near(P::heston_semi_analytic(hp, 100, OptionType::Call), 16.070154917028834, 1e-6, "Heston semi-analytic");
    // Default bumps (finite_difference.cpp:31-38).
    const auto b = G::default_bump_sizes(100, 0.2, 1.0);
    near(b.spot, 1.0, 1e-15, "spot bump");
    near(b.vol, 0.002, 1e-15, "vol bump");
    near(b.rate, 0.01, 1e-15, "rate bump");
    check(mcd::kSobolMaxDimensions == 7, "SOBOL max dimensions stays 7");
}

void engine_error_tests() {
    const auto r = P::monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1000, 42);
    check(std::isnan(r.price) && std::isnan(r.standard_error) && r.path_count == 0,
          "device error reported as {NaN, NaN, 0}");
    check(mcd::last_engine_status() != 0, "engine status set after a failed call");
}

void device_tests() {
    const double bsm = P::black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call);
    const auto e = P::monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1000000, 42);
    check(mcd::last_engine_status() == 0, "engine status OK");
    check(e.path_count == 1000000, "path count");
    near(e.price, bsm, 3.0 * e.standard_error, "European within 3 SE of BSM");
    P::McOptions anti;
    anti.antithetic = true;
    const auto ea = P::monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1000000, 42, anti);
    check(ea.standard_error < e.standard_error, "antithetic lowers the SE");
    near(ea.price, bsm, 3.0 * ea.standard_error, "antithetic within 3 SE");
    // Digital against its closed form.
    const auto d = P::monte_carlo_digital(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                          mcd::DigitalStyle::CashOrNothing, 1.0, 1 << 20, 7);
    near(d.price, P::digital(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, mcd::DigitalStyle::CashOrNothing),
         4.0 * d.standard_error, "digital within 4 SE");
    // Geometric Asian with many dates approaches Kemna-Vorst; CV Asian has a smaller SE.
    const auto g = P::monte_carlo_asian(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, mcd::StrikeStyle::Fixed,
                                        mcd::AverageStyle::Geometric, 252, 1 << 18, 3);
    near(g.price, P::kemna_vorst(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call), 4.0 * g.standard_error + 0.02,
         "geometric Asian near Kemna-Vorst");
    P::McOptions cv;
    cv.control_variate = true;
    const auto a = P::monte_carlo_asian(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, mcd::StrikeStyle::Fixed,
                                        mcd::AverageStyle::Arithmetic, 64, 1 << 18, 3);
    const auto acv = P::monte_carlo_asian(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, mcd::StrikeStyle::Fixed,
                                          mcd::AverageStyle::Arithmetic, 64, 1 << 18, 3, cv);
    check(acv.standard_error < 0.2 * a.standard_error, "control variate cuts the Asian SE");
    near(acv.price, a.price, 4.0 * a.standard_error, "CV Asian consistent");
    // Barrier in + out = vanilla path by path.
    const auto in = P::monte_carlo_barrier(100, 100, 120, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                           mcd::BarrierDirection::Up, mcd::BarrierKnock::In, 0.0, 50, 1 << 18, 9);
    const auto out = P::monte_carlo_barrier(100, 100, 120, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                            mcd::BarrierDirection::Up, mcd::BarrierKnock::Out, 0.0, 50, 1 << 18, 9);
    const auto van = P::monte_carlo_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1 << 18, 9);
    check(in.price > 0.0 && out.price > 0.0, "barrier prices positive");
    near(in.price + out.price, P::black_scholes_merton(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call),
         4.0 * van.standard_error + 0.05, "barrier in + out near the vanilla");
    // Lookback floating call against its continuous closed form (discrete monitoring is lower).
    const auto lb = P::monte_carlo_lookback(100, 0.0, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                            mcd::StrikeStyle::Floating, 252, 1 << 18, 11);
    const double lbc = P::lookback_floating_strike(100, 0.05, 0.0, 0.2, 1.0, OptionType::Call);
    check(lb.price < lbc + 3.0 * lb.standard_error && lb.price > lbc - 1.5, "discrete lookback below continuous");
    // American put: LSM near CRR (the reference's acceptance: 3 SE + Bermudan gap).
    const auto am = P::monte_carlo_lsm_american(100, 100, 0.06, 0.0, 0.2, 1.0, OptionType::Put, 50, 1 << 18, 42);
    near(am.price, P::crr_binomial_american(100, 100, 0.06, 0.0, 0.2, 1.0, 2000, OptionType::Put),
         3.0 * am.standard_error + 0.05, "LSM American put near CRR");
    check(am.policy.coefficients_by_date.size() == 51, "policy has a slot per date");
    const double frozen = P::reprice_against_frozen_policy(100, 100, 0.06, 0.0, 0.2, 1.0, OptionType::Put, 50,
                                                           1 << 18, 42, am.policy);
    near(frozen, am.price, 1e-12 * am.price, "frozen replay reproduces the fitted price");
    // Greeks: FD, LR and pathwise deltas against BSM delta.
    const double h = 1e-4;
    const double bsm_delta = (P::black_scholes_merton(100 + h, 100, 0.05, 0, 0.2, 1, OptionType::Call) -
                              P::black_scholes_merton(100 - h, 100, 0.05, 0, 0.2, 1, OptionType::Call)) / (2 * h);
    const auto fd = G::finite_difference_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1 << 20, 42,
                                                  G::default_bump_sizes(100, 0.2, 1.0));
    near(fd.delta, bsm_delta, 0.01, "FD delta");
    const auto lr = G::likelihood_ratio_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1 << 20, 42);
    near(lr.delta.value, bsm_delta, 4.0 * lr.delta.standard_error, "LR delta");
    check(lr.theta.has_value(), "European LR theta present");
    const auto pw = G::pathwise_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1 << 20, 42);
    near(pw.delta.value, bsm_delta, 4.0 * pw.delta.standard_error, "pathwise delta");
    const auto lb_lr = G::likelihood_ratio_barrier(100, 100, 120, 0.05, 0.0, 0.2, 1.0, OptionType::Call,
                                                   mcd::BarrierDirection::Up, mcd::BarrierKnock::Out, 0.0, 50,
                                                   1 << 16, 11);
    check(!lb_lr.theta.has_value(), "barrier LR omits theta");
    // QMC and Heston.
    near(P::qmc_sobol_european(100, 100, 0.05, 0.0, 0.2, 1.0, OptionType::Call, 1 << 16).price, bsm, 5e-3,
         "QMC European");
    const mcd::models::HestonParams hp{100, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7, 1.0};
    const auto hq = P::heston_qe_european(hp, 100, OptionType::Call, 252, 1 << 20, 42);
    near(hq.price, P::heston_semi_analytic(hp, 100, OptionType::Call), 3.0 * hq.standard_error,
         "Heston QE within 3 SE of semi-analytic");
}

} // namespace

int main(int argc, char** argv) {
    const bool device = argc > 1 && std::strcmp(argv[1], "device") == 0;
    host_tests();
    if (device) device_tests();
    else engine_error_tests();
    std::printf("%d checks, %d failures\n", g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}
