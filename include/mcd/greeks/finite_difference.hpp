// Finite-difference Greeks (reference include/mcd/greeks/finite_difference.hpp:9-33).
#pragma once

#include <cstdint>

#include "mcd/core/types.hpp"

namespace mcd::greeks {

struct EuropeanGreeks {
    double delta;
    double gamma;
    double vega;
    double theta;
    double rho;
};

struct BumpSizes {
    double spot;
    double vol;
    double rate;
    double time;
};

[[nodiscard]] BumpSizes default_bump_sizes(double spot, double vol, double time) noexcept;
[[nodiscard]] EuropeanGreeks finite_difference_european(double spot, double strike, double rate,
                                                          double carry_yield, double vol,
                                                          double time, OptionType type,
                                                          std::uint64_t path_count,
                                                          std::uint64_t seed,
                                                          BumpSizes bumps) noexcept;

} // namespace mcd::greeks
