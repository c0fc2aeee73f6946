// Drop-in counterpart of the reference's include/mcd/core/stats.hpp:21-79: Welford's one-pass
// mean and variance and the Chan-Golub-LeVeque merge, header-only like the reference's. The
// device kernels accumulate with the same update and merge (device_common.hpp).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace mcd {

namespace detail {
inline constexpr std::size_t kDestructiveInterferenceSize = 128; // bytes per padded slot
} // namespace detail

class WelfordAccumulator {
  public:
    void add(double value) noexcept {
        count_ += 1;
        const double before = value - mean_;
        mean_ += before / static_cast<double>(count_);
        m2_ += before * (value - mean_);
    }

    [[nodiscard]] std::uint64_t count() const noexcept { return count_; }
    [[nodiscard]] double mean() const noexcept { return mean_; }

    // sample variance (n - 1); 0 below two samples
    [[nodiscard]] double variance() const noexcept {
        if (count_ < 2) return 0.0;
        return m2_ / static_cast<double>(count_ - 1);
    }

    [[nodiscard]] double standard_error() const noexcept {
        if (count_ == 0) return 0.0;
        return std::sqrt(variance() / static_cast<double>(count_));
    }

    // Folds `other` into this accumulator; an empty side is an exact identity.
    void merge(const WelfordAccumulator& other) noexcept {
        if (other.count_ == 0) return;
        if (count_ == 0) {
            *this = other;
            return;
        }
        const std::uint64_t n = count_ + other.count_;
        const double d = other.mean_ - mean_;
        const double mean = mean_ + d * (static_cast<double>(other.count_) / static_cast<double>(n));
        const double m2 = m2_ + other.m2_ +
                          d * d * static_cast<double>(count_) * static_cast<double>(other.count_) /
                              static_cast<double>(n);
        count_ = n;
        mean_ = mean;
        m2_ = m2;
    }

  private:
    std::uint64_t count_ = 0;
    double mean_ = 0.0;
    double m2_ = 0.0;
};

// One accumulator per cache-line-aligned slot (arrays of them never share a line).
struct alignas(detail::kDestructiveInterferenceSize) PaddedWelford {
    WelfordAccumulator acc;
};

} // namespace mcd
