// HestonParams (reference include/mcd/models/heston.hpp:8-18), same field order.
#pragma once

namespace mcd::models {

struct HestonParams {
    double spot;
    double rate;
    double carry_yield;
    double v0;
    double kappa;
    double theta;
    double xi;
    double rho;
    double time;
};

} // namespace mcd::models
