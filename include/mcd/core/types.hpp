// Drop-in counterpart of the reference's include/mcd/core/types.hpp:12-17: the same enum
// names, underlying type and values, so code written against `mcd` compiles unchanged.
#pragma once

#include <cstdint>
#include <string_view>

namespace mcd {

[[nodiscard]] std::string_view version() noexcept;

enum class OptionType : std::uint8_t { Call, Put };
enum class BarrierDirection : std::uint8_t { Up, Down };
enum class BarrierKnock : std::uint8_t { In, Out };
enum class StrikeStyle : std::uint8_t { Fixed, Floating };
enum class DigitalStyle : std::uint8_t { CashOrNothing, AssetOrNothing };
enum class AverageStyle : std::uint8_t { Arithmetic, Geometric };

} // namespace mcd
