// Engine status for the noexcept C++ API (no reference counterpart). The pricers keep the
// reference's signatures and report a failed device call as a NaN price; this returns the
// calling thread's status of its most recent engine call (an mcdx_status value, 0 = OK).
#pragma once

namespace mcd {

int last_engine_status() noexcept;

} // namespace mcd
