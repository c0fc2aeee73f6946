// Drop-in counterpart of the reference's include/mcd/core/rng.hpp:15-38: Philox4x32-10 and
// the per-path draws, same names, signatures and defaults. Host functions of libmcd.so for
// single draws; bulk generation runs in the device path kernels with the same counters.
#pragma once

#include <array>
#include <cstdint>

namespace mcd {

using PhiloxCounter = std::array<std::uint32_t, 4>;
using PhiloxKey = std::array<std::uint32_t, 2>;

// Philox4x32-10 (Salmon et al. 2011): ten rounds of the 4x32 bijection under `key`.
[[nodiscard]] PhiloxCounter philox4x32_10(PhiloxCounter counter, PhiloxKey key) noexcept;

// key = {seed low word, seed high word}
[[nodiscard]] PhiloxKey make_philox_key(std::uint64_t seed) noexcept;

// counter = {path low word, path high word, draw, stream}; stream 0 feeds the normal
// draws, stream 1 the Brownian-bridge barrier uniforms.
[[nodiscard]] PhiloxCounter make_philox_counter(std::uint64_t path_index,
                                                 std::uint32_t draw_index = 0,
                                                 std::uint32_t stream = 0) noexcept;

// Phi^-1((w0 + 0.5) / 2^32) of Philox word 0 at (seed, path, draw, stream 0).
[[nodiscard]] double standard_normal_variate(std::uint64_t seed, std::uint64_t path_index,
                                              std::uint32_t draw_index = 0) noexcept;

// (w0 + 0.5) / 2^32 of Philox word 0 at (seed, path, draw, stream 1): in (0, 1).
[[nodiscard]] double uniform_variate(std::uint64_t seed, std::uint64_t path_index,
                                      std::uint32_t draw_index) noexcept;

} // namespace mcd
