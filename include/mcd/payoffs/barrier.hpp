// Drop-in counterpart of the reference's include/mcd/payoffs/barrier.hpp:14-38: a discretely
// monitored single barrier, touched when a monitored price reaches it (>= up, <= down); the
// rebate is paid at expiry.
#pragma once

#include "mcd/core/types.hpp"
#include "mcd/payoffs/path_payoff.hpp"

namespace mcd::payoffs {

struct BarrierPayoff {
    double strike = 0.0;
    double barrier = 0.0;
    double rebate = 0.0;
    OptionType type = OptionType::Call;
    BarrierDirection direction = BarrierDirection::Down;
    BarrierKnock knock = BarrierKnock::Out;

    bool breached = false;
    double last_price = 0.0;

    void observe(double price) noexcept {
        if (direction == BarrierDirection::Up ? price >= barrier : price <= barrier) breached = true;
        last_price = price;
    }

    [[nodiscard]] double result() const noexcept {
        const bool alive = knock == BarrierKnock::Out ? !breached : breached;
        if (!alive) return rebate;
        const double v = (type == OptionType::Call ? 1.0 : -1.0) * (last_price - strike);
        return v < 0.0 ? 0.0 : v; // std::max(v, 0.0)
    }
};

static_assert(PathPayoff<BarrierPayoff>);

} // namespace mcd::payoffs
