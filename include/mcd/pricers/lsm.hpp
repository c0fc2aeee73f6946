// Longstaff-Schwartz (reference include/mcd/pricers/lsm.hpp:16-58), on the device.
#pragma once

#include <cstdint>
#include <vector>

#include "mcd/core/types.hpp"

namespace mcd::pricers {

struct LsmPolicy {
    std::vector<std::vector<double>> coefficients_by_date;
    bool exercise_at_inception = false;
};

struct LsmResult {
    double price = 0.0;
    double standard_error = 0.0;
    LsmPolicy policy;
};

[[nodiscard]] LsmResult monte_carlo_lsm_american(double spot, double strike, double rate,
                                                   double carry_yield, double vol, double time,
                                                   OptionType type, int monitoring_points,
                                                   std::uint64_t path_count,
                                                   std::uint64_t seed) noexcept;
[[nodiscard]] double reprice_against_frozen_policy(double spot, double strike, double rate,
                                                     double carry_yield, double vol, double time,
                                                     OptionType type, int monitoring_points,
                                                     std::uint64_t path_count, std::uint64_t seed,
                                                     const LsmPolicy& policy) noexcept;

} // namespace mcd::pricers
