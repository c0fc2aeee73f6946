// Sobol points (reference include/mcd/core/sobol.hpp:14-18). Dimensions 0-6 are the
// reference's; kSobolMaxDimensions stays 7 for drop-in behaviour, the device table extends
// to kSobolExtendedDimensions.
#pragma once

#include <cstdint>

namespace mcd {

constexpr unsigned kSobolMaxDimensions = 7;
constexpr unsigned kSobolExtendedDimensions = 1024;

// x / 2^32 of the Sobol integer at (dimension, index); index 0 -> 0.0. NaN on device error.
[[nodiscard]] double sobol_point(unsigned dimension, std::uint64_t index) noexcept;

} // namespace mcd
