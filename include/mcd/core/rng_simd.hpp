// Drop-in counterpart of the reference's include/mcd/core/rng_simd.hpp: the four-draw batch
// entry point. This library has no NEON path (kHasNeon is false on every host it builds
// for, as the reference's bindings report it); the batch equals four scalar calls bit for bit.
#pragma once

#include <array>
#include <cstdint>

namespace mcd {

inline constexpr bool kHasNeon = false;

// standard_normal_variate(seed, path_index_base + i, draw_index) for i = 0..3.
[[nodiscard]] std::array<double, 4> standard_normal_variate_batch4(
    std::uint64_t seed, std::uint64_t path_index_base, std::uint32_t draw_index = 0) noexcept;

} // namespace mcd
