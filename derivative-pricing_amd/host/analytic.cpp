// Host-side closed forms and lattices of the drop-in API: Black-Scholes-Merton, Black-76,
// Kemna-Vorst, Reiner-Rubinstein barriers, Goldman-Sosin-Gatto lookbacks, digitals,
// forwards (reference src/pricers/analytic.cpp:12-235), CRR (src/pricers/binomial.cpp:9-72),
// Heston's semi-analytic price (src/pricers/heston.cpp:26-82, Gauss-Legendre
// src/core/gauss_legendre.cpp:17-69) and the default bump sizes
// (src/greeks/finite_difference.cpp:24-38). These are deterministic formulas, not Monte
// Carlo; they run on the host by design.
#include <algorithm>
#include <cmath>
#include <complex>
#include <numbers>
#include <vector>

#include "mcd/core/gauss_legendre.hpp"
#include "mcd/core/types.hpp"
#include "mcd/greeks/finite_difference.hpp"
#include "mcd/pricers/analytic.hpp"
#include "mcd/pricers/binomial.hpp"
#include "mcd/pricers/heston.hpp"

namespace mcd {

std::string_view version() noexcept { return "0.1.0"; }

namespace {
double ncdf(double x) { return 0.5 * std::erfc(-x / std::numbers::sqrt2); }
double sgn(OptionType t) { return t == OptionType::Call ? 1.0 : -1.0; }
} // namespace

namespace pricers {

double black_scholes_merton(double s, double k, double r, double q, double v, double t,
                            OptionType type) noexcept {
    const double b = r - q;
    const double st = std::sqrt(t);
    const double d1 = (std::log(s / k) + (b + 0.5 * v * v) * t) / (v * st);
    const double d2 = d1 - v * st;
    const double phi = sgn(type);
    return phi * s * std::exp(-q * t) * ncdf(phi * d1) - phi * k * std::exp(-r * t) * ncdf(phi * d2);
}

double black76(double f, double k, double r, double v, double t, OptionType type) noexcept {
    const double st = std::sqrt(t);
    const double d1 = (std::log(f / k) + 0.5 * v * v * t) / (v * st);
    const double d2 = d1 - v * st;
    const double phi = sgn(type);
    return phi * std::exp(-r * t) * (f * ncdf(phi * d1) - k * ncdf(phi * d2));
}

double kemna_vorst(double s, double k, double r, double q, double v, double t,
                   OptionType type) noexcept {
    const double sigma_g = v / std::numbers::sqrt3;
    const double b_g = 0.5 * (r - q - v * v / 6.0);
    return black_scholes_merton(s, k, r, r - b_g, sigma_g, t, type);
}

double reiner_rubinstein(double s, double k, double h, double r, double q, double v, double t,
                         OptionType type, BarrierDirection direction, BarrierKnock knock,
                         double rebate) noexcept {
    const double b = r - q;
    const double vst = v * std::sqrt(t);
    const double mu = (b - 0.5 * v * v) / (v * v);
    const double lambda = std::sqrt(mu * mu + 2.0 * r / (v * v));
    const double phi = sgn(type);
    const double eta = direction == BarrierDirection::Down ? 1.0 : -1.0;
    const double x1 = std::log(s / k) / vst + (1.0 + mu) * vst;
    const double x2 = std::log(s / h) / vst + (1.0 + mu) * vst;
    const double y1 = std::log(h * h / (s * k)) / vst + (1.0 + mu) * vst;
    const double y2 = std::log(h / s) / vst + (1.0 + mu) * vst;
    const double z = std::log(h / s) / vst + lambda * vst;
    const double dq = std::exp((b - r) * t), dr = std::exp(-r * t);
    const double p1 = std::pow(h / s, 2.0 * (mu + 1.0)), p0 = std::pow(h / s, 2.0 * mu);
    const double A = phi * s * dq * ncdf(phi * x1) - phi * k * dr * ncdf(phi * x1 - phi * vst);
    const double B = phi * s * dq * ncdf(phi * x2) - phi * k * dr * ncdf(phi * x2 - phi * vst);
    const double C = phi * s * dq * p1 * ncdf(eta * y1) - phi * k * dr * p0 * ncdf(eta * y1 - eta * vst);
    const double D = phi * s * dq * p1 * ncdf(eta * y2) - phi * k * dr * p0 * ncdf(eta * y2 - eta * vst);
    const double E = rebate * dr * (ncdf(eta * x2 - eta * vst) - p0 * ncdf(eta * y2 - eta * vst));
    const double F = rebate * (std::pow(h / s, mu + lambda) * ncdf(eta * z) +
                               std::pow(h / s, mu - lambda) * ncdf(eta * z - 2.0 * eta * lambda * vst));
    const bool above = k > h;
    const bool call = type == OptionType::Call, down = direction == BarrierDirection::Down;
    if (knock == BarrierKnock::In) {
        if (down) return call ? (above ? C + E : A - B + D + E) : (above ? B - C + D + E : A + E);
        return call ? (above ? A + E : B - C + D + E) : (above ? A - B + D + E : C + E);
    }
    if (down) return call ? (above ? A - C + F : B - D + F) : (above ? A - B + C - D + F : F);
    return call ? (above ? F : A - B + C - D + F) : (above ? B - D + F : A - C + F);
}

namespace {
// Goldman-Sosin-Gatto lookback pieces (analytic.cpp:146-182).
double lookback_ab(double s, double ref, double r, double q, double v, double t, double eta) {
    const double lambda = 2.0 * (r - q) / (v * v);
    const double sd = v * std::sqrt(t);
    const double d1 = std::log(s / ref) / sd + 0.5 * (lambda + 1.0) * sd;
    const double dq = std::exp(-q * t), dr = std::exp(-r * t);
    return eta * (s * dq * ncdf(eta * d1) - ref * dr * ncdf(eta * (d1 - sd)) -
                  s * dr * (std::pow(s / ref, -lambda) * ncdf(eta * (d1 - lambda * sd)) -
                            dq * ncdf(eta * d1) / dr) / lambda);
}
} // namespace

double lookback_fixed_strike(double s, double k, double r, double q, double v, double t,
                             OptionType type) noexcept {
    const double eta = sgn(type);
    const bool guaranteed = type == OptionType::Call ? k <= s : k >= s;
    if (guaranteed) return lookback_ab(s, s, r, q, v, t, eta) + eta * std::exp(-r * t) * (s - k);
    return lookback_ab(s, k, r, q, v, t, eta);
}

double lookback_floating_strike(double s, double r, double q, double v, double t,
                                OptionType type) noexcept {
    const double eta = sgn(type);
    const double lambda = 2.0 * (r - q) / (v * v);
    const double sd = v * std::sqrt(t);
    const double d1 = 0.5 * (lambda + 1.0) * sd; // log(s/s) = 0
    const double dq = std::exp(-q * t), dr = std::exp(-r * t);
    return eta * ((s * dq * ncdf(eta * d1) - s * dr * ncdf(eta * (d1 - sd))) +
                  s * dr * (ncdf(eta * (-d1 + lambda * sd)) - dq * ncdf(eta * -d1) / dr) / lambda);
}

double digital(double s, double k, double r, double q, double v, double t, OptionType type,
               DigitalStyle style, double cash) noexcept {
    const double st = std::sqrt(t);
    const double d1 = (std::log(s / k) + (r - q + 0.5 * v * v) * t) / (v * st);
    const double phi = sgn(type);
    if (style == DigitalStyle::CashOrNothing) return cash * std::exp(-r * t) * ncdf(phi * (d1 - v * st));
    return s * std::exp(-q * t) * ncdf(phi * d1);
}

double forward_price(double s, double r, double q, double t) noexcept { return s * std::exp((r - q) * t); }
double forward_value(double fn, double f0, double r, double t) noexcept { return (fn - f0) * std::exp(-r * t); }
double futures_mark_to_market(double a, double b) noexcept { return a - b; }

namespace {
struct Lattice {
    double dt, u, d, pi, disc, phi;
};
Lattice lattice(double r, double q, double v, double t, int steps, OptionType type) {
    Lattice l;
    l.dt = t / steps;
    l.u = std::exp(v * std::sqrt(l.dt));
    l.d = 1.0 / l.u;
    l.pi = (std::exp((r - q) * l.dt) - l.d) / (l.u - l.d);
    l.disc = std::exp(-r * l.dt);
    l.phi = sgn(type);
    return l;
}
} // namespace

BinomialResult crr_binomial(double s, double k, double r, double q, double v, double t, int steps,
                            OptionType type) noexcept {
    const Lattice l = lattice(r, q, v, t, steps, type);
    std::vector<double> val(steps + 1);
    for (int i = 0; i <= steps; ++i)
        val[i] = std::max(l.phi * (s * std::pow(l.u, i) * std::pow(l.d, steps - i) - k), 0.0);
    for (int level = steps - 1; level >= 0; --level)
        for (int i = 0; i <= level; ++i) val[i] = l.disc * (l.pi * val[i + 1] + (1.0 - l.pi) * val[i]);
    return BinomialResult{val[0], l.pi, l.u, l.d};
}

double crr_binomial_american(double s, double k, double r, double q, double v, double t, int steps,
                             OptionType type) noexcept {
    const Lattice l = lattice(r, q, v, t, steps, type);
    std::vector<double> val(steps + 1);
    for (int i = 0; i <= steps; ++i)
        val[i] = std::max(l.phi * (s * std::pow(l.u, i) * std::pow(l.d, steps - i) - k), 0.0);
    for (int level = steps - 1; level >= 0; --level)
        for (int i = 0; i <= level; ++i) {
            const double cont = l.disc * (l.pi * val[i + 1] + (1.0 - l.pi) * val[i]);
            const double ex = std::max(l.phi * (s * std::pow(l.u, i) * std::pow(l.d, level - i) - k), 0.0);
            val[i] = std::max(cont, ex);
        }
    return val[0];
}

namespace {
using cd = std::complex<double>;
cd heston_cf(const models::HestonParams& p, cd u) { // Little-Trap form (heston.cpp:26-40)
    const cd i(0.0, 1.0);
    const cd bmr = p.kappa - p.rho * p.xi * i * u;
    const cd d = std::sqrt(bmr * bmr + p.xi * p.xi * (i * u + u * u));
    const cd c = (bmr - d) / (bmr + d);
    const cd e = std::exp(-d * p.time);
    const cd big_c = i * u * (p.rate - p.carry_yield) * p.time +
                     (p.kappa * p.theta / (p.xi * p.xi)) *
                         ((bmr - d) * p.time - 2.0 * std::log((1.0 - c * e) / (1.0 - c)));
    const cd big_d = (bmr - d) / (p.xi * p.xi) * (1.0 - e) / (1.0 - c * e);
    return std::exp(big_c + big_d * p.v0);
}

} // namespace

double heston_semi_analytic(const models::HestonParams& p, double strike, OptionType type) noexcept {
    const cd i(0.0, 1.0);
    const double lm = std::log(strike / p.spot);
    const cd phi_ni = heston_cf(p, cd(0.0, -1.0));
    const GaussLegendreRule rule = gauss_legendre_rule(128); // 128 nodes on [0, 200] (heston.cpp:42-44)
    const std::vector<double>& x = rule.nodes;
    const std::vector<double>& w = rule.weights;
    const double hw = 100.0, mid = 100.0;
    double s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < 128; ++k) {
        const double u = mid + hw * x[k];
        s1 += w[k] * (std::exp(-i * u * lm) * (heston_cf(p, cd(u, -1.0)) / phi_ni) / (i * u)).real();
        s2 += w[k] * (std::exp(-i * u * lm) * heston_cf(p, cd(u, 0.0)) / (i * u)).real();
    }
    const double p1 = 0.5 + hw * s1 / std::numbers::pi, p2 = 0.5 + hw * s2 / std::numbers::pi;
    const double dq = std::exp(-p.carry_yield * p.time), dr = std::exp(-p.rate * p.time);
    const double call = p.spot * dq * p1 - strike * dr * p2;
    return type == OptionType::Call ? call : call - p.spot * dq + strike * dr;
}

} // namespace pricers

namespace greeks {
BumpSizes default_bump_sizes(double spot, double vol, double time) noexcept {
    return BumpSizes{std::max(spot * 0.01, 1e-4), std::max(vol * 0.01, 1e-4), 0.01,
                     std::max(time * 0.01, 1e-4)};
}
} // namespace greeks

} // namespace mcd
