// Basic types of the suite harness (src/common/RPTypes.hpp:49-107): FP64
// data, 64-bit indices and a long-double checksum.
#pragma once

#include <cstddef>
#include <cstdint>

namespace rajaperf {

using RepIndex_type = volatile int;  // RPTypes.hpp:49
using Index_type = std::int64_t;     // RPTypes.hpp:59 (RAJA::Index_type)
using Size_type = std::size_t;       // RPTypes.hpp:71
using Checksum_type = long double;   // RPTypes.hpp:93
using Real_type = double;            // RPTypes.hpp:107
using Real_ptr = Real_type*;

}  // namespace rajaperf
