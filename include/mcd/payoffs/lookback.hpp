// Drop-in counterpart of the reference's include/mcd/payoffs/lookback.hpp:13-57: running
// extrema seeded at the initial spot.
#pragma once

#include "mcd/core/types.hpp"
#include "mcd/payoffs/path_payoff.hpp"

namespace mcd::payoffs {

// max(M - K, 0) for a call, max(K - m, 0) for a put (M, m: running max, min)
struct LookbackFixedPayoff {
    double strike;
    OptionType type;
    double running_max;
    double running_min;

    LookbackFixedPayoff(double initial_spot, double strike_, OptionType type_) noexcept
        : strike(strike_), type(type_), running_max(initial_spot), running_min(initial_spot) {}

    void observe(double price) noexcept {
        if (price > running_max) running_max = price;
        if (price < running_min) running_min = price;
    }

    [[nodiscard]] double result() const noexcept {
        const double v = type == OptionType::Call ? running_max - strike : strike - running_min;
        return v < 0.0 ? 0.0 : v; // std::max(v, 0.0)
    }
};

static_assert(PathPayoff<LookbackFixedPayoff>);

// S_T - m for a call, M - S_T for a put (never negative)
struct LookbackFloatingPayoff {
    OptionType type;
    double last_price;
    double running_max;
    double running_min;

    LookbackFloatingPayoff(double initial_spot, OptionType type_) noexcept
        : type(type_), last_price(initial_spot), running_max(initial_spot),
          running_min(initial_spot) {}

    void observe(double price) noexcept {
        if (price > running_max) running_max = price;
        if (price < running_min) running_min = price;
        last_price = price;
    }

    [[nodiscard]] double result() const noexcept {
        return type == OptionType::Call ? last_price - running_min : running_max - last_price;
    }
};

static_assert(PathPayoff<LookbackFloatingPayoff>);

} // namespace mcd::payoffs
