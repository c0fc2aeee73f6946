// Pathwise Greeks (reference include/mcd/greeks/pathwise.hpp:14-60).
#pragma once

#include <cstdint>

#include "mcd/core/types.hpp"

namespace mcd::greeks {

struct PathwiseGreeksResult {
    double value = 0.0;
    double standard_error = 0.0;
};

struct PathwiseGreeks {
    PathwiseGreeksResult delta, vega, rho;
};

[[nodiscard]] PathwiseGreeks pathwise_european(double spot, double strike, double rate,
                                                double carry_yield, double vol, double time,
                                                OptionType type, std::uint64_t path_count,
                                                std::uint64_t seed) noexcept;
[[nodiscard]] PathwiseGreeks pathwise_asian(double spot, double strike, double rate,
                                             double carry_yield, double vol, double time,
                                             OptionType type, StrikeStyle strike_style,
                                             AverageStyle average_style, int monitoring_points,
                                             std::uint64_t path_count,
                                             std::uint64_t seed) noexcept;

} // namespace mcd::greeks
