// Device execution of renormalisation plans (host/renorm.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>

#include "../host/renorm.hpp"
#include "hip_util.hpp"

namespace t3 {

// A result arena left on the device (besides its host copy).
using DevArena = std::shared_ptr<DevBuf<double>>;

// Runs an append plan on the device: the old operator arena (host data of
// in.old, or `dev_old` when it is already on the device) is uploaded, the
// new arena computed, all-zero blocks dropped (src/sparseblocks.c:73-107),
// and the result returned with host data; `keep` receives its device copy.
struct RenormWork;
ROperators run_append(const AppendPlan& plan, const ROperators& old, hipStream_t s,
                      RenormWork* work = nullptr, const double* dev_old = nullptr,
                      DevArena* keep = nullptr);

// Work of one device execution: algorithmic flops and bytes (inputs read
// once, outputs written once) and the summed kernel time.
struct RenormWork {
    double flops = 0, bytes = 0, kernel_ms = 0;
};

// Runs an update_rOperators_physical plan on the device: every contribution
// is two GEMMs (the cheaper association of A*O*B is chosen per block pair),
// the first into a device workspace, the second summed over the block's
// contributions in plan order; all-zero blocks are then dropped.
// Runs an update_rOperators_branching plan: per contribution the tensor is
// contracted with the two old operator blocks (two GEMM stages into a device
// workspace, processed in chunks of destination blocks), the adjoint block
// closes the contraction in a fixed-order sum per unique-operator block, and
// sum_unique_rOperators with zero-block removal finishes on the device.
// Old operators already on the device come as dev0 / dev1 / dev_pops
// (else from their host data); `keep` receives the result's device copy.
ROperators run_bupdate(const BUpdatePlan& plan, const SiteTensor& T, const ROperators& ops0,
                       const ROperators& ops1, hipStream_t s, RenormWork* work,
                       const double* dev0 = nullptr, const double* dev1 = nullptr,
                       DevArena* keep = nullptr);

ROperators run_pupdate(const PUpdatePlan& plan, const ROperators& pops, const SiteTensor& T,
                       hipStream_t s, RenormWork* work, const double* dev_pops = nullptr,
                       DevArena* keep = nullptr);

}  // namespace t3
