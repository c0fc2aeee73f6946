// Likelihood-ratio Greeks (reference include/mcd/greeks/likelihood_ratio.hpp:15-60).
#pragma once

#include <cstdint>
#include <optional>

#include "mcd/core/types.hpp"

namespace mcd::greeks {

struct LrGreeksResult {
    double value = 0.0;
    double standard_error = 0.0;
};

struct LrGreeks {
    LrGreeksResult delta, gamma, vega, rho;
    std::optional<LrGreeksResult> theta; // not computed for barriers
};

[[nodiscard]] LrGreeks likelihood_ratio_european(double spot, double strike, double rate,
                                                   double carry_yield, double vol, double time,
                                                   OptionType type, std::uint64_t path_count,
                                                   std::uint64_t seed) noexcept;
[[nodiscard]] LrGreeks likelihood_ratio_digital(double spot, double strike, double rate,
                                                  double carry_yield, double vol, double time,
                                                  OptionType type, DigitalStyle style,
                                                  double cash_amount, std::uint64_t path_count,
                                                  std::uint64_t seed) noexcept;
[[nodiscard]] LrGreeks likelihood_ratio_barrier(double spot, double strike, double barrier,
                                                  double rate, double carry_yield, double vol,
                                                  double time, OptionType type,
                                                  BarrierDirection direction, BarrierKnock knock,
                                                  double rebate, int monitoring_points,
                                                  std::uint64_t path_count,
                                                  std::uint64_t seed) noexcept;

} // namespace mcd::greeks
