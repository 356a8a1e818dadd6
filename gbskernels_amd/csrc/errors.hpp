// errors.hpp -- return codes of the C ABI. Zero is success; HIP runtime
// failures are returned as their hipError_t value (all below 10000); the
// library's own argument errors are the GBSK_ERR_* codes of include/gbsk.h.
#pragma once

#include "../../include/gbsk.h"

namespace gbsk {
constexpr int GBSK_OK = 0;
}  // namespace gbsk
