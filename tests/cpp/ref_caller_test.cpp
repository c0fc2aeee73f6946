// A caller written against the reference's public headers the way its own benchmark and core
// tests are: it includes what bench/mc_bench.cpp:1-4, tests/rng_test.cpp, tests/stats_test.cpp,
// tests/linalg_test.cpp, tests/gauss_legendre_test.cpp and the payoff and model headers
// include, compiles against include/mcd/** and links libmcd.so. The checks restate those
// tests' assertions (Random123 known answers, the Newton inverse-CDF oracle and its corner
// bound, moments and Kolmogorov-Smirnov at 1e6 draws, Welford against two passes, the
// textbook least-squares systems, Gauss-Legendre exactness) plus the payoff structs'
// definitions and the bridge's law.
//
// Usage: ref_caller_test host            -- host functions only
//        ref_caller_test device          -- also a Monte Carlo pricing through the headers
//        ref_caller_test draws s p d ... -- print standard_normal_variate and uniform_variate
//                                           of each (seed, path, draw) as hex floats
//        ref_caller_test bridge n s p    -- print brownian_bridge_gbm_path's spots
//        ref_caller_test lsq r c s       -- print householder_least_squares's coefficients
#include "mcd/core/rng.hpp"
#include "mcd/core/rng_simd.hpp"
#include "mcd/core/stats.hpp"
#include "mcd/pricers/monte_carlo.hpp"

#include "mcd/core/gauss_legendre.hpp"
#include "mcd/core/linalg.hpp"
#include "mcd/core/normal.hpp"
#include "mcd/models/brownian_bridge_path.hpp"
#include "mcd/models/gbm.hpp"
#include "mcd/payoffs/asian.hpp"
#include "mcd/payoffs/barrier.hpp"
#include "mcd/payoffs/digital.hpp"
#include "mcd/payoffs/european.hpp"
#include "mcd/payoffs/lookback.hpp"
#include "mcd/payoffs/path_payoff.hpp"
#include "mcd/pricers/analytic.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <numbers>
#include <string>
#include <vector>

namespace {

int g_fail = 0, g_checks = 0;

void expect(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        ++g_fail;
        std::printf("FAIL: %s\n", what.c_str());
    }
}

void near(double got, double want, double tol, const char* what) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: got %.17g want %.17g tol %.3g", what, got, want, tol);
    expect(std::fabs(got - want) <= tol, buf);
}

double newton_inverse(double u) {
    double z = 0.0;
    for (int i = 0; i < 100; ++i) {
        const double step = (mcd::standard_normal_cdf(z) - u) / mcd::standard_normal_pdf(z);
        z -= step;
        if (std::fabs(step) < 1e-15) break;
    }
    return z;
}

void rng_checks() {
    // Random123 known answers (reference tests/rng_test.cpp:17-39)
    const mcd::PhiloxCounter kat[3][3] = {
        {{0, 0, 0, 0}, {0, 0, 0, 0}, {0x6627E8D5u, 0xE169C58Du, 0xBC57AC4Cu, 0x9B00DBD8u}},
        {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, {0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0},
         {0x408F276Du, 0x41C83B0Eu, 0xA20BC7C6u, 0x6D5451FDu}},
        {{0x243F6A88u, 0x85A308D3u, 0x13198A2Eu, 0x03707344u}, {0xA4093822u, 0x299F31D0u, 0, 0},
         {0xD16CFE09u, 0x94FDCCEBu, 0x5001E420u, 0x24126EA1u}}};
    for (const auto& v : kat) {
        const mcd::PhiloxKey key{v[1][0], v[1][1]};
        expect(mcd::philox4x32_10(v[0], key) == v[2], "Philox4x32-10 known answer");
    }
    expect(mcd::make_philox_key(0x299F31D0A4093822ull) == mcd::PhiloxKey{0xA4093822u, 0x299F31D0u},
           "key layout");
    expect(mcd::make_philox_counter(0x85A308D3243F6A88ull, 7) ==
               mcd::PhiloxCounter{0x243F6A88u, 0x85A308D3u, 7u, 0u},
           "counter layout and defaults");
    // inverse normal against Newton on Phi (rng_test.cpp:76-117)
    double worst = 0.0;
    for (double p = 1e-9; p < 0.5; p *= 3.16227766017)
        for (double u : {p, 1.0 - p}) worst = std::max(worst, std::fabs(mcd::inverse_standard_normal_cdf(u) - newton_inverse(u)));
    near(worst, 0.0, 1e-9, "inverse normal vs Newton on [1e-9, 1 - 1e-9]");
    for (double p : {1e-12, 1e-11, 1e-10})
        for (double u : {p, 1.0 - p}) {
            const double z = mcd::inverse_standard_normal_cdf(u);
            const double floor = std::numeric_limits<double>::epsilon() / mcd::standard_normal_pdf(z);
            expect(std::fabs(z - newton_inverse(u)) < 50.0 * floor, "inverse normal corner bound");
        }
    for (double z : {-6.0, -3.0, -1.0, -0.1, 0.0, 0.1, 1.0, 3.0, 6.0})
        near(mcd::inverse_standard_normal_cdf(mcd::standard_normal_cdf(z)), z, 1e-8, "round trip");
    // moments and Kolmogorov-Smirnov of 1e6 draws (rng_test.cpp:131-188)
    const std::uint64_t n = 1000000;
    std::vector<double> x(n);
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (std::uint64_t i = 0; i < n; ++i) {
        const double z = mcd::standard_normal_variate(1, i);
        s1 += z;
        s2 += z * z;
        s3 += z * z * z;
        s4 += z * z * z * z;
        x[i] = mcd::standard_normal_variate(2, i);
    }
    const double nd = (double)n, m = s1 / nd, var = s2 / nd - m * m;
    const double skew = (s3 / nd - 3 * m * s2 / nd + 2 * m * m * m) / std::pow(var, 1.5);
    const double kurt = (s4 / nd - 4 * m * s3 / nd + 6 * m * m * s2 / nd - 3 * m * m * m * m) / (var * var) - 3.0;
    near(m, 0.0, 5.0 / std::sqrt(nd), "mean");
    near(var, 1.0, 5.0 * std::sqrt(2.0 / nd), "variance");
    near(skew, 0.0, 5.0 * std::sqrt(6.0 / nd), "skewness");
    near(kurt, 0.0, 5.0 * std::sqrt(24.0 / nd), "excess kurtosis");
    std::sort(x.begin(), x.end());
    double d = 0.0;
    for (std::uint64_t i = 0; i < n; ++i) {
        const double f = mcd::standard_normal_cdf(x[i]);
        d = std::max({d, std::fabs((double)(i + 1) / nd - f), std::fabs(f - (double)i / nd)});
    }
    expect(d < 1.63 / std::sqrt(nd), "Kolmogorov-Smirnov at the 1% level");
    // uniforms lie in (0, 1) and the batch equals four scalar draws
    for (std::uint64_t i = 0; i < 1000; ++i) {
        const double u = mcd::uniform_variate(5, i, 3);
        expect(u > 0.0 && u < 1.0, "uniform in (0, 1)");
    }
    const auto b4 = mcd::standard_normal_variate_batch4(42, 1000, 9);
    for (int i = 0; i < 4; ++i) expect(b4[i] == mcd::standard_normal_variate(42, 1000 + i, 9), "batch4 = scalar");
    expect(!mcd::kHasNeon, "no NEON path");
}

void stats_checks() {
    // Welford against the two-pass formulas on 1e6 draws (stats_test.cpp:9-33)
    const std::uint64_t n = 1000000;
    std::vector<double> v(n);
    mcd::WelfordAccumulator acc;
    for (std::uint64_t i = 0; i < n; ++i) acc.add(v[i] = mcd::standard_normal_variate(99, i));
    double sum = 0;
    for (double x : v) sum += x;
    const double mean = sum / (double)n;
    double ss = 0;
    for (double x : v) ss += (x - mean) * (x - mean);
    near(acc.mean(), mean, 1e-12, "Welford mean");
    near(acc.variance(), ss / (double)(n - 1), 1e-12, "Welford variance");
    expect(acc.count() == n, "Welford count");
    near(acc.standard_error(), std::sqrt(acc.variance() / (double)n), 1e-15, "standard error");
    // merges: empty is an identity; two disjoint halves equal the whole (stats_test.cpp:52-104)
    mcd::WelfordAccumulator a, b, e;
    for (std::uint64_t i = 0; i < 300000; ++i) a.add(v[i]);
    for (std::uint64_t i = 300000; i < n; ++i) b.add(v[i]);
    mcd::WelfordAccumulator a2 = a;
    a2.merge(e);
    expect(a2.count() == a.count() && a2.mean() == a.mean() && a2.variance() == a.variance(), "merge empty right");
    e.merge(a);
    expect(e.count() == a.count() && e.mean() == a.mean() && e.variance() == a.variance(), "merge empty left");
    a.merge(b);
    near(a.mean(), mean, 1e-9, "merged mean");
    near(a.variance(), ss / (double)(n - 1), 1e-9, "merged variance");
    expect(sizeof(mcd::PaddedWelford) % mcd::detail::kDestructiveInterferenceSize == 0, "padded slot");
}

void linalg_checks() {
    // textbook systems (linalg_test.cpp:39-55)
    std::vector<double> a = {1, 1, 1, 2, 1, 3};
    auto beta = mcd::householder_least_squares(a, 3, 2, {1, 2, 2});
    near(beta[0], 2.0 / 3.0, 1e-9, "line fit b0");
    near(beta[1], 0.5, 1e-9, "line fit b1");
    std::vector<double> sq = {2, 1, 1, 1, 3, 2, 1, 0, 0};
    beta = mcd::householder_least_squares(sq, 3, 3, {4, 5, 6});
    near(beta[0], 6.0, 1e-9, "square b0");
    near(beta[1], 15.0, 1e-9, "square b1");
    near(beta[2], -23.0, 1e-9, "square b2");
    // an ill-conditioned cubic Vandermonde system recovers its coefficients
    const double xs[4] = {1.0, 1.0001, 1.0002, 1.0003}, bt[4] = {2.0, -3.0, 5.0, -1.0};
    std::vector<double> v, y;
    for (double x : xs) {
        double p = 1.0, yy = 0.0;
        for (int k = 0; k < 4; ++k) {
            v.push_back(p);
            yy += bt[k] * p;
            p *= x;
        }
        y.push_back(yy);
    }
    beta = mcd::householder_least_squares(v, 4, 4, y);
    double resid = 0.0;
    for (int i = 0; i < 4; ++i) {
        double p = 1.0, fit = 0.0;
        for (int k = 0; k < 4; ++k) {
            fit += beta[k] * p;
            p *= xs[i];
        }
        resid = std::max(resid, std::fabs(fit - y[i]));
    }
    near(resid, 0.0, 1e-9, "Vandermonde residual");
}

void gauss_legendre_checks() {
    // exact for degree <= 2 order - 1 (gauss_legendre_test.cpp:12-22)
    for (int order : {2, 4, 8, 16, 32, 64})
        for (int k = 0; k <= 2 * order - 1; ++k) {
            const double got = mcd::gauss_legendre_integrate([k](double x) { return std::pow(x, k); }, -1.0, 1.0, order);
            near(got, k % 2 == 0 ? 2.0 / (k + 1) : 0.0, 1e-10, "Gauss-Legendre exactness");
        }
    const auto rule = mcd::gauss_legendre_rule(32);
    double sum = 0.0;
    for (double w : rule.weights) {
        expect(w > 0.0, "positive weight");
        sum += w;
    }
    near(sum, 2.0, 1e-10, "weights sum");
    expect(std::is_sorted(rule.nodes.begin(), rule.nodes.end()), "ascending nodes");
    near(mcd::gauss_legendre_integrate([](double x) { return x * x * x; }, 0.0, 2.0, 4), 4.0, 1e-10, "int x^3");
    near(mcd::gauss_legendre_integrate([](double x) { return std::sin(x); }, 0.0, std::numbers::pi / 2, 64), 1.0, 1e-12, "int sin");
}

template <mcd::payoffs::PathPayoff P>
double run_path(P p, const std::vector<double>& prices) {
    for (double s : prices) p.observe(s);
    return p.result();
}

void payoff_and_model_checks() {
    using mcd::OptionType;
    namespace pf = mcd::payoffs;
    static_assert(pf::Payoff<pf::EuropeanPayoff> && pf::Payoff<pf::DigitalPayoff>);
    static_assert(pf::PathPayoff<pf::ArithmeticAsianPayoff> && pf::PathPayoff<pf::BarrierPayoff>);
    expect(pf::EuropeanPayoff{100.0, OptionType::Call}(112.0) == 12.0, "European call");
    expect(pf::EuropeanPayoff{100.0, OptionType::Put}(112.0) == 0.0, "European put");
    expect(pf::DigitalPayoff{100.0, OptionType::Call, mcd::DigitalStyle::CashOrNothing, 2.5}(101.0) == 2.5, "cash digital");
    expect(pf::DigitalPayoff{100.0, OptionType::Call, mcd::DigitalStyle::AssetOrNothing, 1.0}(101.0) == 101.0, "asset digital");
    expect(pf::DigitalPayoff{100.0, OptionType::Call}(100.0) == 0.0, "digital at the strike");
    const std::vector<double> path = {95.0, 104.0, 121.0, 110.0};
    near(run_path(pf::ArithmeticAsianPayoff{100.0, OptionType::Call}, path), 7.5, 1e-15, "arithmetic Asian");
    near(run_path(pf::ArithmeticAsianPayoff{100.0, OptionType::Put, mcd::StrikeStyle::Floating}, path), 0.0, 0.0, "floating Asian put");
    const double geo = std::exp((std::log(95.0) + std::log(104.0) + std::log(121.0) + std::log(110.0)) / 4.0);
    near(run_path(pf::GeometricAsianPayoff{100.0, OptionType::Call}, path), geo - 100.0, 1e-12, "geometric Asian");
    pf::BarrierPayoff uo{100.0, 120.0, 1.5, OptionType::Call, mcd::BarrierDirection::Up, mcd::BarrierKnock::Out};
    expect(run_path(uo, path) == 1.5, "up-and-out touched: rebate");
    pf::BarrierPayoff ui = uo;
    ui.knock = mcd::BarrierKnock::In;
    expect(run_path(ui, path) == 10.0, "up-and-in touched: vanilla");
    pf::BarrierPayoff touch_eq = uo;
    touch_eq.barrier = 121.0;
    expect(run_path(touch_eq, path) == 1.5, "touch at equality (>=)");
    expect(run_path(pf::LookbackFixedPayoff(100.0, 105.0, OptionType::Call), path) == 16.0, "fixed lookback call");
    expect(run_path(pf::LookbackFixedPayoff(100.0, 105.0, OptionType::Put), path) == 10.0, "fixed lookback put");
    expect(run_path(pf::LookbackFloatingPayoff(100.0, OptionType::Call), path) == 15.0, "floating lookback call");
    expect(run_path(pf::LookbackFloatingPayoff(100.0, OptionType::Put), path) == 11.0, "floating lookback put");
    // exact GBM step
    const mcd::models::GbmParams g{100.0, 0.05, 0.01, 0.2, 2.0};
    near(mcd::models::gbm_terminal_spot(g, 0.3),
         100.0 * std::exp((0.05 - 0.01 - 0.02) * 2.0 + 0.2 * std::sqrt(2.0) * 0.3), 1e-12, "GBM terminal");
    // bridge: W(T) = sqrt(T) z0 (the terminal spot equals the exact step with z0), and with
    // one date the bridge is that step
    std::vector<double> z(13);
    for (int i = 0; i < 13; ++i) z[i] = mcd::standard_normal_variate(3, 7, (std::uint32_t)i);
    const auto spots = mcd::models::brownian_bridge_gbm_path(g, 13, z);
    near(spots.back(), mcd::models::gbm_terminal_spot(g, z[0]), 1e-12, "bridge terminal");
    const auto one = mcd::models::brownian_bridge_gbm_path(g, 1, z);
    near(one[0], mcd::models::gbm_terminal_spot(g, z[0]), 1e-12, "one-date bridge");
    // the bridge's law: Var W(t_i) = t_i over many paths (each normal enters once)
    const int m = 8, npaths = 200000;
    std::vector<double> sumsq(m, 0.0);
    const mcd::models::GbmParams unit{1.0, 0.0, 0.0, 1.0, 1.0};
    for (int p = 0; p < npaths; ++p) {
        std::vector<double> zz(m);
        for (int i = 0; i < m; ++i) zz[i] = mcd::standard_normal_variate(11, (std::uint64_t)p, (std::uint32_t)i);
        const auto s = mcd::models::brownian_bridge_gbm_path(unit, m, zz);
        for (int i = 0; i < m; ++i) {
            const double w = std::log(s[i]) + 0.5 * (i + 1) / (double)m; // W(t_i)
            sumsq[i] += w * w;
        }
    }
    for (int i = 0; i < m; ++i) {
        const double t = (i + 1) / (double)m;
        near(sumsq[i] / npaths, t, 5.0 * t * std::sqrt(2.0 / npaths), "bridge variance");
    }
}

void device_checks() {
    // the reference benchmark's call (bench/mc_bench.cpp:17-24) against Black-Scholes-Merton
    const auto r = mcd::pricers::monte_carlo_european(100.0, 100.0, 0.05, 0.01, 0.25, 1.0,
                                                      mcd::OptionType::Call, 1000000, 0);
    const double bsm = mcd::pricers::black_scholes_merton(100.0, 100.0, 0.05, 0.01, 0.25, 1.0,
                                                          mcd::OptionType::Call);
    expect(r.path_count == 1000000, "path count");
    near(r.price, bsm, 3.0 * r.standard_error, "monte_carlo_european vs BSM");
}

} // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "host";
    if (mode == "draws") {
        for (int i = 2; i + 2 < argc; i += 3) {
            const std::uint64_t seed = std::strtoull(argv[i], nullptr, 0);
            const std::uint64_t path = std::strtoull(argv[i + 1], nullptr, 0);
            const auto draw = (std::uint32_t)std::strtoul(argv[i + 2], nullptr, 0);
            std::printf("%a %a\n", mcd::standard_normal_variate(seed, path, draw),
                        mcd::uniform_variate(seed, path, draw));
        }
        return 0;
    }
    if (mode == "bridge" && argc == 5) { // n seed path: spots from draws 0..n-1 of (seed, path)
        const int n = std::atoi(argv[2]);
        const std::uint64_t seed = std::strtoull(argv[3], nullptr, 0), path = std::strtoull(argv[4], nullptr, 0);
        std::vector<double> z((std::size_t)n);
        for (int i = 0; i < n; ++i) z[(std::size_t)i] = mcd::standard_normal_variate(seed, path, (std::uint32_t)i);
        for (double s : mcd::models::brownian_bridge_gbm_path({100.0, 0.05, 0.01, 0.25, 1.5}, n, z))
            std::printf("%a\n", s);
        return 0;
    }
    if (mode == "lsq" && argc == 5) { // rows cols seed: A and y from draws; prints beta
        const int rows = std::atoi(argv[2]), cols = std::atoi(argv[3]);
        const std::uint64_t seed = std::strtoull(argv[4], nullptr, 0);
        std::vector<double> a((std::size_t)rows * cols), y((std::size_t)rows);
        for (int i = 0; i < rows; ++i) {
            for (int j = 0; j < cols; ++j)
                a[(std::size_t)i * cols + j] = mcd::standard_normal_variate(seed, (std::uint64_t)i, (std::uint32_t)j);
            y[(std::size_t)i] = mcd::standard_normal_variate(seed, (std::uint64_t)i, 99u);
        }
        for (double b : mcd::householder_least_squares(a, rows, cols, y)) std::printf("%a\n", b);
        return 0;
    }
    rng_checks();
    stats_checks();
    linalg_checks();
    gauss_legendre_checks();
    payoff_and_model_checks();
    if (mode == "device") device_checks();
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
