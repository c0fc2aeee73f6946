// GBM Monte Carlo pricers (reference include/mcd/pricers/monte_carlo.hpp:9-80), backed by
// the MI355X engine through include/mcdx.h. Same names, signatures and defaults. On a
// device error the result is {NaN, NaN, 0} (the functions stay noexcept).
#pragma once

#include <cstdint>

#include "mcd/core/types.hpp"

namespace mcd::pricers {

struct McResult {
    double price;
    double standard_error;
    std::uint64_t path_count;
};

struct McOptions {
    bool antithetic = false;
    bool control_variate = false;
    bool brownian_bridge = false;
    unsigned num_threads = 1; // accepted; the device schedule does not depend on it
};

[[nodiscard]] McResult monte_carlo_european(double spot, double strike, double rate,
                                             double carry_yield, double vol, double time,
                                             OptionType type, std::uint64_t path_count,
                                             std::uint64_t seed) noexcept;
[[nodiscard]] McResult monte_carlo_european(double spot, double strike, double rate,
                                             double carry_yield, double vol, double time,
                                             OptionType type, std::uint64_t path_count,
                                             std::uint64_t seed, McOptions options) noexcept;
[[nodiscard]] McResult monte_carlo_digital(double spot, double strike, double rate,
                                            double carry_yield, double vol, double time,
                                            OptionType type, DigitalStyle style,
                                            double cash_amount, std::uint64_t path_count,
                                            std::uint64_t seed, McOptions options = {}) noexcept;
[[nodiscard]] McResult monte_carlo_asian(double spot, double strike, double rate,
                                          double carry_yield, double vol, double time,
                                          OptionType type, StrikeStyle strike_style,
                                          AverageStyle average_style, int monitoring_points,
                                          std::uint64_t path_count, std::uint64_t seed,
                                          McOptions options = {}) noexcept;
[[nodiscard]] McResult monte_carlo_barrier(double spot, double strike, double barrier,
                                            double rate, double carry_yield, double vol,
                                            double time, OptionType type,
                                            BarrierDirection direction, BarrierKnock knock,
                                            double rebate, int monitoring_points,
                                            std::uint64_t path_count, std::uint64_t seed,
                                            McOptions options = {}) noexcept;
[[nodiscard]] McResult monte_carlo_lookback(double spot, double strike, double rate,
                                             double carry_yield, double vol, double time,
                                             OptionType type, StrikeStyle style,
                                             int monitoring_points, std::uint64_t path_count,
                                             std::uint64_t seed, McOptions options = {}) noexcept;

} // namespace mcd::pricers
