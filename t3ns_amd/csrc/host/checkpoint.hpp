// Checkpoints in the reference's HDF5 layout (T3NScalc.h5,
// src/io_to_disk.c:446-487): /network, /bookkeeper, /hamiltonian, /T3NS
// and /rOps.  The HDF5 library is loaded at run time (libhdf5.so.103, or
// the path in T3NS_HDF5_LIB), so the product library does not depend on
// it at link time.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "context.hpp"

namespace t3 {

// One site tensor as the reference stores it (struct siteTensor with its
// sparseblocks, include/siteTensor.h:73-95).
struct SiteBlocks {
    std::vector<int> sites;
    std::vector<int64_t> qnumbers;    // [nrblocks * sites.size()]
    std::vector<int64_t> beginblock;  // [nrblocks + 1]
    std::vector<double> tel;
    int nrblocks() const { return (int)beginblock.size() - 1; }
};

struct CheckpointData {
    Network netw;
    SymInfo sym;
    HamData ham;
    std::vector<Symsecs> v_symsecs, p_symsecs;
    std::vector<SiteBlocks> tensors;  // one per network site
    int max_symmetries = MAX_SYMMETRIES;
    int nr_rops_groups = 0;           // rOperator_i groups present (not read)
};

// Writes the context's network, bookkeeper and Hamiltonian and the given
// site tensors (one per site).  The /rOps group holds nrOps and no
// operators: a restored calculation rebuilds its operators from the site
// tensors.
void write_checkpoint(const std::string& path, const Context& c,
                      const std::vector<SiteBlocks>& tensors);

// Reads a checkpoint written here or by the reference (its /rOps
// operators are counted, not read).
CheckpointData read_checkpoint(const std::string& path);

// A context with the checkpoint's network, symmetries, Hamiltonian and
// bond sectors (in place of prepare_bookkeeper).
std::unique_ptr<Context> context_from_checkpoint(const CheckpointData& d);

}  // namespace t3
