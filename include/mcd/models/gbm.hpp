// Drop-in counterpart of the reference's include/mcd/models/gbm.hpp:5-16.
#pragma once

namespace mcd::models {

struct GbmParams {
    double spot;
    double rate;
    double carry_yield;
    double vol;
    double time;
};

// The exact GBM solution at params.time for one standard normal z:
// S0 exp((r - q - vol^2 / 2) T + vol sqrt(T) z).
[[nodiscard]] double gbm_terminal_spot(const GbmParams& params, double standard_normal) noexcept;

} // namespace mcd::models
