// Device execution of the site-tensor plans (host/sitetensor.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../host/sitetensor.hpp"
#include "hip_util.hpp"
#include "renorm_dev.hpp"

namespace t3 {

// Contracts the one-site tensors into the multisite tensor: per block, one
// in-place contraction per internal bond (strided GEMM tiles, all blocks of
// one bond in one launch), then the reorder of the inner site's legs into
// site order where needed.  Returns the multisite elements (host); blocks
// without all one-site blocks are zero.
std::vector<double> run_makesite(const MakesitePlan& plan, const SiteTensor* const* sites,
                                 hipStream_t s, RenormWork* work);

}  // namespace t3
