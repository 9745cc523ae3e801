// Effective-Hamiltonian dispatch plan: which (output block, input block,
// MPO combination) triples contribute, with which operator blocks and which
// symmetry prefactor.  This is the table the reference builds in
// init_Heffdata + the first matvec (src/Heff.c:108-484, 695-914, 1002-1047,
// 1110-1221) and caches in `struct secondrun` (include/Heff.h:36-51).
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"
#include "sym.hpp"
#include "tensors.hpp"

namespace t3 {

// Everything init_Heffdata reads, passed explicitly instead of via the
// reference's globals.
struct HeffInput {
    SymInfo sym;
    const SiteTensor* site = nullptr;
    bool site_is_phys[STEPSPECS_MSITES] = {false, false, false, false};
    const Symsecs* symarr[STEPSPECS_MSITES][3] = {};
    const Symsecs* mpo = nullptr;
    const ROperators* ops[3] = {nullptr, nullptr, nullptr};
    int rops_on_site[3] = {-1, -1, -1};
    const InstructionSet* iset = nullptr;  // merge set, sorted and grouped
    std::vector<int> hermitian;            // [nrhss]
    std::vector<unsigned char> singlet;    // [nrhss^3], index a + b*H + c*H*H
};

struct PlanEntry {     // one MPO combination for one block pair
    int mpo_group;     // index into iset.MPOc
    int sbops[3];      // operator block index (within the hss group)
    double prefactor;  // symmetry prefactor
};

struct PlanPair {  // one (new block, old block) pair
    int newsb, oldsb;
    int bestorder;
    std::vector<PlanEntry> entries;
};

struct HeffPlan {
    int isdmrg = 0;
    int posB = 0;
    int map[3] = {0, 1, 2};
    int nrblocks = 0;
    int nr_qnB = 0;
    std::vector<std::array<int, 3>> dims;  // leg dims per block (op order)
    std::vector<int64_t> beginblock;       // block offsets in the vector
    std::vector<PlanPair> pairs;           // grouped by newsb, newsb ascending
    std::vector<int> pair_begin;           // [nrblocks+1] into pairs
    const InstructionSet* iset = nullptr;
    // Algorithmic flops of one matvec under the reference's contraction
    // order: sum over executed GEMMs of 2*M*N*K*L.
    double flops = 0;
    int64_t nterms = 0;
};

HeffPlan build_heff_plan(const HeffInput& in);

// Contraction order minimising M*N*K work (src/Heff.c:629-666).
int heff_best_order(const int dold[3], const int dnew[3], int isdmrg);
// Flops of one term under an order.
double heff_term_flops(const int dold[3], const int dnew[3], const int map[3], int order, int isdmrg);

// Diagonal contributions (src/Heff.c:1064-1108, 1268-1312): for every block,
// the MPO groups of its own (block, block) pair with operator blocks and
// symmetry prefactor.  Zero-prefactor groups are kept, as in the reference.
struct DiagEntry {
    int sb;
    int mpo_group;
    int sbops[3];
    double prefactor;
};
std::vector<DiagEntry> build_heff_diag_plan(const HeffInput& in, const HeffPlan& plan);

}  // namespace t3
