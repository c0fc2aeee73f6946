// Sobol QMC pricers (reference include/mcd/pricers/qmc.hpp:9-34), on the device.
#pragma once

#include <cstdint>

#include "mcd/core/types.hpp"

namespace mcd::pricers {

struct QmcResult {
    double price = 0.0;
};

[[nodiscard]] QmcResult qmc_sobol_european(double spot, double strike, double rate,
                                            double carry_yield, double vol, double time,
                                            OptionType type, std::uint64_t path_count) noexcept;
// monitoring_points in [1, 7] as the reference; the device table extends further through
// mcdx_qmc_sobol_asian_extended (include/mcdx.h).
[[nodiscard]] QmcResult qmc_sobol_asian(double spot, double strike, double rate,
                                         double carry_yield, double vol, double time,
                                         OptionType type, StrikeStyle strike_style,
                                         int monitoring_points, std::uint64_t path_count) noexcept;

} // namespace mcd::pricers
