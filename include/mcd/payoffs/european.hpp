// Drop-in counterpart of the reference's include/mcd/payoffs/european.hpp:10-23.
#pragma once

#include <concepts>

#include "mcd/core/types.hpp"

namespace mcd::payoffs {

// A payoff of the terminal spot alone.
template <typename P>
concept Payoff = requires(const P& p, double terminal_spot) {
    { p(terminal_spot) } -> std::convertible_to<double>;
};

struct EuropeanPayoff {
    double strike = 0.0;
    OptionType type = OptionType::Call;

    // max(phi (S_T - K), 0), phi = +1 call, -1 put
    [[nodiscard]] double operator()(double terminal_spot) const noexcept {
        const double v = (type == OptionType::Call ? 1.0 : -1.0) * (terminal_spot - strike);
        return v < 0.0 ? 0.0 : v; // std::max(v, 0.0)
    }
};

static_assert(Payoff<EuropeanPayoff>);

} // namespace mcd::payoffs
