// hip_error.hpp -- C ABI status to C++ exception, shared by the C++ headers over brp_hip.h.
// The reference's backend throws std::runtime_error with its message; the C ABI returns a
// BRP_E_* class and keeps the message for brp_last_error().
#pragma once
#include <stdexcept>

#include "brp_hip.h"

namespace brp::hip_detail {

inline void check(int rc) {
    if (rc != BRP_OK) throw std::runtime_error(brp_last_error());
}

}  // namespace brp::hip_detail
