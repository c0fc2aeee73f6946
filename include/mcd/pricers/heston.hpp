// Heston pricers (reference include/mcd/pricers/heston.hpp:10-33): QE Monte Carlo on the
// device, the semi-analytic characteristic-function price on the host.
#pragma once

#include <cstdint>

#include "mcd/core/types.hpp"
#include "mcd/models/heston.hpp"

namespace mcd::pricers {

struct HestonMcResult {
    double price = 0.0;
    double standard_error = 0.0;
    std::uint64_t path_count = 0;
};

[[nodiscard]] HestonMcResult heston_qe_european(const models::HestonParams& params,
                                                 double strike, OptionType type,
                                                 int monitoring_points, std::uint64_t path_count,
                                                 std::uint64_t seed) noexcept;
[[nodiscard]] double heston_semi_analytic(const models::HestonParams& params, double strike,
                                           OptionType type) noexcept;

} // namespace mcd::pricers
