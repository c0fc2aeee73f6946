// Drop-in counterpart of the reference's include/mcd/core/linalg.hpp: least squares by
// Householder QR (a host function; the device LSM regression is a tall-skinny QR of the same
// problem, lsm.hip).
#pragma once

#include <vector>

namespace mcd {

// min ||A beta - y||_2 for row-major A (rows x cols, rows >= cols), A overwritten; returns
// beta (cols). A column that is zero from the diagonal down gets no reflection and a zero
// coefficient.
[[nodiscard]] std::vector<double> householder_least_squares(std::vector<double>& a, int rows,
                                                              int cols,
                                                              std::vector<double> y) noexcept;

} // namespace mcd
