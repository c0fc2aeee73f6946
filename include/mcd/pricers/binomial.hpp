// CRR lattice (reference include/mcd/pricers/binomial.hpp:7-28), host-side.
#pragma once

#include "mcd/core/types.hpp"

namespace mcd::pricers {

struct BinomialResult {
    double price;
    double risk_neutral_probability;
    double up_factor;
    double down_factor;
};

[[nodiscard]] BinomialResult crr_binomial(double spot, double strike, double rate,
                                           double carry_yield, double vol, double time, int steps,
                                           OptionType type) noexcept;
[[nodiscard]] double crr_binomial_american(double spot, double strike, double rate,
                                            double carry_yield, double vol, double time, int steps,
                                            OptionType type) noexcept;

} // namespace mcd::pricers
