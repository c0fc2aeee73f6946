// Drop-in counterpart of the reference's include/mcd/payoffs/path_payoff.hpp:12-16: a payoff
// that observes each monitored price and settles once, holding O(1) state. The device path
// kernels evaluate these payoffs with the same definitions (gbm_paths.hpp).
#pragma once

#include <concepts>

namespace mcd::payoffs {

template <typename P>
concept PathPayoff = requires(P& p, double price) {
    { p.observe(price) } -> std::same_as<void>;
    { p.result() } -> std::convertible_to<double>;
};

} // namespace mcd::payoffs
