// time_call (reference include/mcd/core/timing.hpp:19-26): wall time of one call.
#pragma once

#include <chrono>
#include <utility>

namespace mcd {

template <typename T>
struct TimingResult {
    T value;
    double elapsed_seconds;
};

template <typename F>
auto time_call(F&& f) -> TimingResult<decltype(f())> {
    const auto start = std::chrono::steady_clock::now();
    auto value = std::forward<F>(f)();
    const auto stop = std::chrono::steady_clock::now();
    return {std::move(value), std::chrono::duration<double>(stop - start).count()};
}

} // namespace mcd
