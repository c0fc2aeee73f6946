// Drop-in counterpart of the reference's include/mcd/models/brownian_bridge_path.hpp:24-26:
// Brownian-bridge path construction (a host function; the device QMC kernels run the same
// construction from a host-built program, sobol.hip).
#pragma once

#include <span>
#include <vector>

#include "mcd/models/gbm.hpp"

namespace mcd::models {

// Spots at t_i = i T / n, i = 1..n (chronological), from n normals taken in construction
// order: normals[0] sets W(T) = sqrt(T) normals[0], and each later normal fills the midpoint
// of the next interval in breadth-first bisection order from its conditional law. The caller
// guarantees n >= 1 and normals.size() >= n.
[[nodiscard]] std::vector<double> brownian_bridge_gbm_path(const GbmParams& params,
                                                             int monitoring_points,
                                                             std::span<const double> normals) noexcept;

} // namespace mcd::models
