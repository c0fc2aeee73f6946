// Host primitives of the drop-in C++ API: Philox4x32-10 and the per-path draws
// (reference src/core/rng.cpp:9-74), the normal distribution (src/core/normal.cpp:50-82),
// Householder least squares (src/core/linalg.cpp:7-71), Gauss-Legendre rules
// (src/core/gauss_legendre.cpp:17-69), the exact GBM step (src/models/gbm.cpp:7-12) and the
// Brownian-bridge path (src/models/brownian_bridge_path.cpp:20-77). These are scalar building
// blocks for callers of the C++ headers; every pricer runs on the device. Built without FP
// contraction, so the FP64 expressions round as the reference's do.
#include <cmath>
#include <numbers>
#include <span>
#include <vector>

#include "acklam.hpp"
#include "mcd/core/gauss_legendre.hpp"
#include "mcd/core/linalg.hpp"
#include "mcd/core/normal.hpp"
#include "mcd/core/rng.hpp"
#include "mcd/core/rng_simd.hpp"
#include "mcd/models/brownian_bridge_path.hpp"
#include "mcd/models/gbm.hpp"

namespace mcd {

namespace {
constexpr std::uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u; // round multipliers
constexpr std::uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u; // key schedule (Weyl) increments

std::uint32_t philox_word0(std::uint64_t seed, std::uint64_t path, std::uint32_t draw,
                           std::uint32_t stream) {
    return philox4x32_10(make_philox_counter(path, draw, stream), make_philox_key(seed))[0];
}

double open01(std::uint32_t bits) { return (static_cast<double>(bits) + 0.5) / 4294967296.0; }
} // namespace

PhiloxCounter philox4x32_10(PhiloxCounter x, PhiloxKey k) noexcept {
    for (int round = 0; round < 10; ++round) {
        const std::uint64_t a = static_cast<std::uint64_t>(kM0) * x[0];
        const std::uint64_t b = static_cast<std::uint64_t>(kM1) * x[2];
        const std::uint32_t ahi = static_cast<std::uint32_t>(a >> 32), alo = static_cast<std::uint32_t>(a);
        const std::uint32_t bhi = static_cast<std::uint32_t>(b >> 32), blo = static_cast<std::uint32_t>(b);
        x = PhiloxCounter{bhi ^ x[1] ^ k[0], blo, ahi ^ x[3] ^ k[1], alo};
        k[0] += kW0;
        k[1] += kW1;
    }
    return x;
}

PhiloxKey make_philox_key(std::uint64_t seed) noexcept {
    return PhiloxKey{static_cast<std::uint32_t>(seed), static_cast<std::uint32_t>(seed >> 32)};
}

PhiloxCounter make_philox_counter(std::uint64_t path_index, std::uint32_t draw_index,
                                  std::uint32_t stream) noexcept {
    return PhiloxCounter{static_cast<std::uint32_t>(path_index),
                         static_cast<std::uint32_t>(path_index >> 32), draw_index, stream};
}

double standard_normal_variate(std::uint64_t seed, std::uint64_t path_index,
                               std::uint32_t draw_index) noexcept {
    return inverse_standard_normal_cdf(open01(philox_word0(seed, path_index, draw_index, 0)));
}

double uniform_variate(std::uint64_t seed, std::uint64_t path_index,
                       std::uint32_t draw_index) noexcept {
    return open01(philox_word0(seed, path_index, draw_index, 1));
}

std::array<double, 4> standard_normal_variate_batch4(std::uint64_t seed,
                                                     std::uint64_t path_index_base,
                                                     std::uint32_t draw_index) noexcept {
    std::array<double, 4> out;
    for (std::uint64_t i = 0; i < 4; ++i)
        out[i] = standard_normal_variate(seed, path_index_base + i, draw_index);
    return out;
}

double standard_normal_pdf(double x) noexcept {
    return std::exp(-0.5 * x * x) / std::sqrt(2.0 * std::numbers::pi);
}

double standard_normal_cdf(double x) noexcept {
    return 0.5 * std::erfc(-x / std::numbers::sqrt2);
}

double inverse_standard_normal_cdf(double u) noexcept {
    MCDX_ACKLAM_CENTRAL(double, static_cast<double>)
    MCDX_ACKLAM_TAIL(double, static_cast<double>)
    constexpr double kLow = 0.02425, kHigh = 1.0 - kLow; // central / tail break
    double z;
    if (u < kLow || kHigh < u) {
        const bool lower = u < kLow;
        const double t = std::sqrt(-2.0 * std::log(lower ? u : 1.0 - u));
        const double r = (((((C1 * t + C2) * t + C3) * t + C4) * t + C5) * t + C6) /
                         ((((D1 * t + D2) * t + D3) * t + D4) * t + 1.0);
        z = lower ? r : -r;
    } else {
        const double q = u - 0.5;
        const double r = q * q;
        z = (((((A1 * r + A2) * r + A3) * r + A4) * r + A5) * r + A6) * q /
            (((((B1 * r + B2) * r + B3) * r + B4) * r + B5) * r + 1.0);
    }
    // one Halley step on Phi(z) = u
    const double e = standard_normal_cdf(z) - u;
    const double r = e * std::sqrt(2.0 * std::numbers::pi) * std::exp(0.5 * z * z);
    return z - r / (1.0 + 0.5 * z * r);
}

std::vector<double> householder_least_squares(std::vector<double>& a, int rows, int cols,
                                              std::vector<double> y) noexcept {
    auto at = [&](int i, int j) -> double& { return a[static_cast<std::size_t>(i) * cols + j]; };
    std::vector<double> v(static_cast<std::size_t>(rows > 0 ? rows : 1));
    for (int k = 0; k < cols; ++k) {
        // reflector v = x - alpha e1 of column k below the diagonal, alpha = -sign(x1) |x|
        double nsq = 0.0;
        for (int i = k; i < rows; ++i) nsq += at(i, k) * at(i, k);
        const double norm = std::sqrt(nsq);
        if (norm == 0.0) continue; // a zero column needs no reflection
        const double alpha = at(k, k) >= 0.0 ? -norm : norm;
        v[0] = at(k, k) - alpha;
        for (int i = k + 1; i < rows; ++i) v[static_cast<std::size_t>(i - k)] = at(i, k);
        double vv = 0.0;
        for (int i = 0; i < rows - k; ++i) vv += v[static_cast<std::size_t>(i)] * v[static_cast<std::size_t>(i)];
        if (vv == 0.0) continue;
        // H = I - 2 v v^T / (v^T v) applied to columns k.. and to y
        for (int j = k; j < cols; ++j) {
            double d = 0.0;
            for (int i = k; i < rows; ++i) d += v[static_cast<std::size_t>(i - k)] * at(i, j);
            const double f = 2.0 * d / vv;
            for (int i = k; i < rows; ++i) at(i, j) -= f * v[static_cast<std::size_t>(i - k)];
        }
        double d = 0.0;
        for (int i = k; i < rows; ++i) d += v[static_cast<std::size_t>(i - k)] * y[static_cast<std::size_t>(i)];
        const double f = 2.0 * d / vv;
        for (int i = k; i < rows; ++i) y[static_cast<std::size_t>(i)] -= f * v[static_cast<std::size_t>(i - k)];
    }
    // back substitution on R (zero diagonal: zero coefficient)
    std::vector<double> beta(static_cast<std::size_t>(cols), 0.0);
    for (int i = cols - 1; i >= 0; --i) {
        double s = y[static_cast<std::size_t>(i)];
        for (int j = i + 1; j < cols; ++j) s -= at(i, j) * beta[static_cast<std::size_t>(j)];
        const double diag = at(i, i);
        beta[static_cast<std::size_t>(i)] = diag != 0.0 ? s / diag : 0.0;
    }
    return beta;
}

GaussLegendreRule gauss_legendre_rule(int order) {
    GaussLegendreRule rule;
    const int n = order;
    rule.nodes.assign(static_cast<std::size_t>(n), 0.0);
    rule.weights.assign(static_cast<std::size_t>(n), 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        // Newton on P_n from the asymptotic guess for the i-th largest root
        double z = std::cos(std::numbers::pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z; // Bonnet's recurrence up to P_n(z)
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double step = p1 / dp;
            z -= step;
            if (std::abs(step) < 1e-15) break;
        }
        rule.nodes[static_cast<std::size_t>(i)] = -z;
        rule.nodes[static_cast<std::size_t>(n - 1 - i)] = z;
        const double w = 2.0 / ((1.0 - z * z) * dp * dp);
        rule.weights[static_cast<std::size_t>(i)] = w;
        rule.weights[static_cast<std::size_t>(n - 1 - i)] = w;
    }
    return rule;
}

namespace models {

double gbm_terminal_spot(const GbmParams& p, double z) noexcept {
    const double drift = (p.rate - p.carry_yield - 0.5 * p.vol * p.vol) * p.time;
    return p.spot * std::exp(drift + p.vol * std::sqrt(p.time) * z);
}

std::vector<double> brownian_bridge_gbm_path(const GbmParams& p, int n,
                                             std::span<const double> normals) noexcept {
    std::vector<double> t(static_cast<std::size_t>(n) + 1), w(static_cast<std::size_t>(n), 0.0);
    for (int i = 0; i <= n; ++i) t[static_cast<std::size_t>(i)] = static_cast<double>(i) * p.time / static_cast<double>(n);
    w[static_cast<std::size_t>(n - 1)] = std::sqrt(t[static_cast<std::size_t>(n)]) * normals[0];
    // Breadth-first bisection of (left, right) index intervals, left = -1 meaning t = 0: W at
    // the midpoint given its ends is normal with the linear interpolation of the ends as mean
    // and variance (t_m - t_l)(t_r - t_m) / (t_r - t_l). Normals are used in that order.
    std::vector<std::pair<int, int>> fifo{{-1, n - 1}};
    std::size_t head = 0;
    int used = 1;
    while (head < fifo.size() && used < n) {
        const auto [l, r] = fifo[head++];
        if (r - l <= 1) continue;
        const int m = (l + r) / 2;
        const double tl = l < 0 ? 0.0 : t[static_cast<std::size_t>(l + 1)];
        const double wl = l < 0 ? 0.0 : w[static_cast<std::size_t>(l)];
        const double tr = t[static_cast<std::size_t>(r + 1)], wr = w[static_cast<std::size_t>(r)];
        const double tm = t[static_cast<std::size_t>(m + 1)];
        const double mean = wl + (tm - tl) / (tr - tl) * (wr - wl);
        const double var = (tm - tl) * (tr - tm) / (tr - tl);
        w[static_cast<std::size_t>(m)] = mean + std::sqrt(var) * normals[static_cast<std::size_t>(used++)];
        fifo.emplace_back(l, m);
        fifo.emplace_back(m, r);
    }
    const double drift_rate = p.rate - p.carry_yield - 0.5 * p.vol * p.vol;
    std::vector<double> spots(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i)
        spots[static_cast<std::size_t>(i)] =
            p.spot * std::exp(drift_rate * t[static_cast<std::size_t>(i + 1)] + p.vol * w[static_cast<std::size_t>(i)]);
    return spots;
}

} // namespace models
} // namespace mcd
