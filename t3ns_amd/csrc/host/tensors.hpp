// Block-sparse site tensors and renormalised operators: metadata layout and
// the structure builders.  Layout follows include/siteTensor.h:73-95 and
// include/rOperators.h:91-123, with 64-bit block offsets (the reference uses
// int offsets, include/sparseblocks.h:48).  Element data lives in one flat
// arena per object; `tel` is the host copy.
#pragma once
#include <cstdint>
#include <vector>

#include "context.hpp"

namespace t3 {

struct SiteTensor {
    int nrsites = 0;
    int sites[STEPSPECS_MSITES] = {-1, -1, -1, -1};
    int nrblocks = 0;
    std::vector<QN> qnumbers;         // nrblocks * nrsites
    std::vector<int64_t> beginblock;  // nrblocks + 1
    std::vector<double> tel;          // host data (may be empty)
    int64_t size() const { return beginblock.empty() ? 0 : beginblock[nrblocks]; }
    int64_t block_size(int b) const { return beginblock[b + 1] - beginblock[b]; }
};

struct ROperators {
    int bond = -1;
    int is_left = -1;
    int P_operator = 0;
    int nrhss = -1;
    std::vector<int> begin_blocks_of_hss;  // nrhss + 1
    std::vector<QN> qnumbers;              // nblocks * (P ? 3 : 1)
    int nrops = 0;
    std::vector<int> hss_of_ops;
    // Per operator: block offsets relative to the operator's start
    // (nblocks_of_op + 1 entries) and the operator's start in the arena.
    std::vector<std::vector<int64_t>> beginblock;
    std::vector<int64_t> op_offset;  // nrops + 1
    std::vector<double> tel;         // host arena (may be empty)
    // Per-operator caller data when `tel` is not copied (the Heff setup
    // uploads straight from it; valid only while the caller's arrays are).
    std::vector<const double*> tel_src;

    int couplings() const { return P_operator ? 3 : 1; }
    int nblocks_hss(int hss) const { return begin_blocks_of_hss[hss + 1] - begin_blocks_of_hss[hss]; }
    const QN* qnumbers_hss(int hss) const {
        return qnumbers.data() + (size_t)begin_blocks_of_hss[hss] * couplings();
    }
    int nblocks_op(int op) const { return nblocks_hss(hss_of_ops[op]); }
    int64_t total_size() const { return op_offset.empty() ? 0 : op_offset[nrops]; }
    // Size of block `b` of operator `op` (0 for an empty block).
    int64_t block_size(int op, int b) const { return beginblock[op][b + 1] - beginblock[op][b]; }
    void finalize_offsets();
};

// Block structure of a one-site tensor from the valid sector triples
// (src/siteTensor_init.c:64-108).  Data is zero-initialised.
SiteTensor init_1siteTensor(const Context& c, int site);
// The same over given sectors of the site's three bonds.
SiteTensor init_1siteTensor(const Context& c, int site, const Symsecs* const ss[3]);

// Structure of renormalised operators at a bond (src/rOperators_init.c:281-507).
// Returns per-hss block offset arrays in `tmpbb`.
ROperators init_rOperators_structure(const Context& c, int bond, int is_left, bool P_operator,
                                     std::vector<std::vector<int64_t>>* tmpbb);
// Unit operator at a border bond (src/rOperators_init.c:59-154).
ROperators vacuum_rOperators(const Context& c, int bond, int is_left);
int rOperators_site_to_attach(const Context& c, const ROperators& r);

// Multisite tensor structure (src/siteTensor_init.c:144-663): computes the
// internal symmetry sectors, installs them into the context and returns the
// zero-initialised multisite tensor.  `info` receives what is needed to fill
// its data from the one-site tensors.
struct MultisiteInfo {
    int nr_internal = 0;
    int internals[3] = {-1, -1, -1};
    int innersite = -1;
    int outersites[3] = {-1, -1, -1};
    int connected[3][2];
    int bonds[STEPSPECS_MSITES][3];
    Symsecs ssarr[STEPSPECS_MSITES][3];      // new (internal bonds replaced)
    Symsecs ssarr_old[STEPSPECS_MSITES][3];  // before
};
SiteTensor make_multisite_structure(Context& c, const int* sitelist, int nr, MultisiteInfo* info);

// Internal bonds of a multisite tensor, in the reference's order
// (src/siteTensor_misc.c:186-214).
std::vector<int> internal_bonds(const Context& c, const SiteTensor& t);

}  // namespace t3
