// Device execution of renormalisation plans (host/renorm.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "../host/renorm.hpp"
#include "hip_util.hpp"

namespace t3 {

// Runs an append plan on the device: the old operator arena (host data of
// in.old) is uploaded, the new arena computed, all-zero blocks dropped
// (src/sparseblocks.c:73-107), and the result returned with host data.
ROperators run_append(const AppendPlan& plan, const ROperators& old, hipStream_t s);

}  // namespace t3
