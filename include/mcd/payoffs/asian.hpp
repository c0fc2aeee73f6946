// Drop-in counterpart of the reference's include/mcd/payoffs/asian.hpp:14-65: averages over
// the monitored dates only (not t = 0); fixed strike compares the average with K, floating
// strike the last price with the average.
#pragma once

#include <cmath>

#include "mcd/core/types.hpp"
#include "mcd/payoffs/path_payoff.hpp"

namespace mcd::payoffs {

namespace detail {
inline double asian_settle(double average, double last, double strike, OptionType type,
                           StrikeStyle style) noexcept {
    const double ref = style == StrikeStyle::Fixed ? average - strike : last - average;
    const double v = (type == OptionType::Call ? 1.0 : -1.0) * ref;
    return v < 0.0 ? 0.0 : v; // std::max(v, 0.0)
}
} // namespace detail

struct ArithmeticAsianPayoff {
    double strike = 0.0;
    OptionType type = OptionType::Call;
    StrikeStyle style = StrikeStyle::Fixed;

    double sum = 0.0;
    double last_price = 0.0;
    unsigned count = 0;

    void observe(double price) noexcept {
        sum += price;
        last_price = price;
        count += 1;
    }

    [[nodiscard]] double result() const noexcept {
        return detail::asian_settle(sum / static_cast<double>(count), last_price, strike, type,
                                    style);
    }
};

static_assert(PathPayoff<ArithmeticAsianPayoff>);

// The geometric average exp(mean log S); also the arithmetic Asian's control variate.
struct GeometricAsianPayoff {
    double strike = 0.0;
    OptionType type = OptionType::Call;
    StrikeStyle style = StrikeStyle::Fixed;

    double log_sum = 0.0;
    double last_price = 0.0;
    unsigned count = 0;

    void observe(double price) noexcept {
        log_sum += std::log(price);
        last_price = price;
        count += 1;
    }

    [[nodiscard]] double result() const noexcept {
        return detail::asian_settle(std::exp(log_sum / static_cast<double>(count)), last_price,
                                    strike, type, style);
    }
};

static_assert(PathPayoff<GeometricAsianPayoff>);

} // namespace mcd::payoffs
