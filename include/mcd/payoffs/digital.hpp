// Drop-in counterpart of the reference's include/mcd/payoffs/digital.hpp:10-24.
#pragma once

#include "mcd/core/types.hpp"
#include "mcd/payoffs/european.hpp"

namespace mcd::payoffs {

struct DigitalPayoff {
    double strike = 0.0;
    OptionType type = OptionType::Call;
    DigitalStyle style = DigitalStyle::CashOrNothing;
    double cash_amount = 1.0;

    // strictly in the money: cash_amount (cash-or-nothing) or S_T (asset-or-nothing); else 0
    [[nodiscard]] double operator()(double terminal_spot) const noexcept {
        const double v = (type == OptionType::Call ? 1.0 : -1.0) * (terminal_spot - strike);
        if (!(v > 0.0)) return 0.0;
        return style == DigitalStyle::CashOrNothing ? cash_amount : terminal_spot;
    }
};

static_assert(Payoff<DigitalPayoff>);

} // namespace mcd::payoffs
