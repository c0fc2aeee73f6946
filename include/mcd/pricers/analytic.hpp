// Closed-form pricers (reference include/mcd/pricers/analytic.hpp:8-56), host-side.
#pragma once

#include "mcd/core/types.hpp"

namespace mcd::pricers {

[[nodiscard]] double black_scholes_merton(double spot, double strike, double rate,
                                           double carry_yield, double vol, double time,
                                           OptionType type) noexcept;
[[nodiscard]] double black76(double forward, double strike, double rate, double vol, double time,
                              OptionType type) noexcept;
[[nodiscard]] double kemna_vorst(double spot, double strike, double rate, double carry_yield,
                                  double vol, double time, OptionType type) noexcept;
[[nodiscard]] double reiner_rubinstein(double spot, double strike, double barrier, double rate,
                                        double carry_yield, double vol, double time,
                                        OptionType type, BarrierDirection direction,
                                        BarrierKnock knock, double rebate = 0.0) noexcept;
[[nodiscard]] double lookback_fixed_strike(double spot, double strike, double rate,
                                            double carry_yield, double vol, double time,
                                            OptionType type) noexcept;
[[nodiscard]] double lookback_floating_strike(double spot, double rate, double carry_yield,
                                               double vol, double time, OptionType type) noexcept;
[[nodiscard]] double digital(double spot, double strike, double rate, double carry_yield,
                              double vol, double time, OptionType type, DigitalStyle style,
                              double cash_amount = 1.0) noexcept;
[[nodiscard]] double forward_price(double spot, double rate, double carry_yield,
                                    double time) noexcept;
[[nodiscard]] double forward_value(double forward_now, double forward_at_inception, double rate,
                                    double time_to_expiry) noexcept;
[[nodiscard]] double futures_mark_to_market(double futures_price_new,
                                             double futures_price_old) noexcept;

} // namespace mcd::pricers
