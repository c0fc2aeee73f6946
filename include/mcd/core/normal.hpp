// Drop-in counterpart of the reference's include/mcd/core/normal.hpp:6-14.
#pragma once

namespace mcd {

// phi(x) = exp(-x^2 / 2) / sqrt(2 pi)
[[nodiscard]] double standard_normal_pdf(double x) noexcept;

// Phi(x) = erfc(-x / sqrt 2) / 2
[[nodiscard]] double standard_normal_cdf(double x) noexcept;

// Phi^-1(u), u in (0, 1): Acklam's rational approximation refined by one Halley step, the
// reference's FP64 arithmetic (the device FP64 mode evaluates the same expressions).
[[nodiscard]] double inverse_standard_normal_cdf(double u) noexcept;

} // namespace mcd
