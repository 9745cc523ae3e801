// Renormalisation plans: which blocks of which operators combine, with which
// symmetry prefactors, when a physical site is appended to renormalised
// operators (src/rOperators_pUpdate.c:393-673, src/rOperators_init.c:
// 185-265).  The plan is executed on the device (dev/renorm_dev.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"
#include "sym.hpp"
#include "tensors.hpp"

namespace t3 {

// Everything rOperators_append_phys reads (explicit instead of globals).
struct AppendInput {
    SymInfo sym;
    int is_left = 0;
    Symsecs site_ss[3], new_ss[3], old_ss[3], mpo;
    std::vector<Instruction> instr;  // sorted pUpdate instructions
    std::vector<int> hss_of_new;
    std::vector<int> siteop_hss;
    int physdim = 0;
    std::vector<double> siteop_el;  // (op*physdim + bra)*physdim + ket
    std::vector<int> tprod_begin, tprods;
    int nrhss = 0;
    std::vector<int> begin_blocks_of_hss;
    std::vector<QN> qnumbers;                   // 3 per block
    std::vector<std::vector<int64_t>> hss_bb;  // per hss block offsets
    const ROperators* old = nullptr;
};

// Destination block of the new operators and the old blocks summed into
// it: dst = sum_t pref_t * (factor_t * src_t), in instruction order
// (the reference first scales into the unique operator, then sums).
struct AppendGroup {
    int64_t dst;
    int64_t n;
    int task_begin, task_end;
    int op, block;  // new operator and its block (within its MPO sector)
};
struct AppendTask {
    int64_t src;
    double factor;  // prefactor_pAppend x site-operator element
    double pref;    // instruction prefactor
};
struct AppendPlan {
    ROperators nr;  // structure of the new operators, all blocks present
    std::vector<AppendGroup> groups;
    std::vector<AppendTask> tasks;
};

AppendPlan plan_append(const AppendInput& in);

// Everything update_rOperators_physical reads.
struct PUpdateInput {
    SymInfo sym;
    int is_left = 0;
    Symsecs site_ss[3], iss, mpo;
    const SiteTensor* T = nullptr;  // one-site tensor (qnumbers over site_ss)
    int nrhss = 0;
    std::vector<int> begin_blocks_of_hss;
    std::vector<QN> qnumbers;
    std::vector<std::vector<int64_t>> hss_bb;
    const ROperators* pops = nullptr;  // physical operators on the internal bond
};
// One contribution O'[dst] += pref * Th^(T) * O[src] * T^(T) of one operator.
struct PUpdateTask {
    int64_t dst, src, th, t;  // offsets: new arena, old arena, tensor, tensor
    int M0, M1, N0, N1;
    double pref;
};
struct PUpdateGroup {
    int task_begin, task_end;
    int op, block;
    int64_t dst, n;
};
struct PUpdatePlan {
    ROperators nr;
    std::vector<PUpdateGroup> groups;
    std::vector<PUpdateTask> tasks;
};
PUpdatePlan plan_pupdate(const PUpdateInput& in);

// Everything update_rOperators_branching reads.
struct BUpdateInput {
    SymInfo sym;
    int update_case = -1;      // bond of the branching site carrying the new operators
    Symsecs site_ss[3], mpo;   // the site's bonds (bra = ket sectors) and the MPO
    const SiteTensor* T = nullptr;
    std::vector<Instruction> instr;  // sorted bUpdate instructions
    std::vector<int> hss_of_new;
    std::vector<int> tprod_begin, tprods;  // per new MPO sector: (hss of op 0, hss of op 1) pairs
    int nrhss = 0;
    std::vector<int> begin_blocks_of_hss;
    std::vector<QN> qnumbers;
    std::vector<std::vector<int64_t>> hss_bb;
    const ROperators* ops[2] = {nullptr, nullptr};
};
// One contribution U[dst] += pref * A . (O1 x O2) . T of one unique operator
// block: A and T are the adjoint and ket blocks of the site tensor, O1 and O2
// blocks of the two old operator sets.  d: ket dims, e: bra dims per bond.
struct BUpdateTask {
    int64_t t, a, o1, o2;  // offsets: tensor, tensor, ops[0] arena, ops[1] arena
    int d[3], e[3];
    double pref;
};
// All contributions of one unique-operator block, in the reference's order.
struct BUpdateDest {
    int64_t dst;  // offset in the unique-operator arena
    int M, N;     // bra x ket dims of the block
    int task_begin, task_end;
};
struct BUpdatePlan {
    int n_leg = -1, x_leg = -1, y_leg = -1;  // bonds of the new operators, ops[0], ops[1]
    ROperators uniq;  // unique operators (all blocks present)
    std::vector<BUpdateDest> dests;
    std::vector<BUpdateTask> tasks;
    // sum_unique_rOperators: new operator blocks from unique blocks
    ROperators nr;
    std::vector<AppendGroup> groups;
    std::vector<AppendTask> stasks;  // src: offset in the unique arena
};
BUpdatePlan plan_bupdate(const BUpdateInput& in);

// Drops the blocks flagged empty: returns the compacted per-operator block
// offsets and, for every kept block, (old offset, new offset, size).
struct Compaction {
    std::vector<std::vector<int64_t>> beginblock;
    std::vector<int64_t> op_offset;
    std::vector<int64_t> src, dst, n;
};
Compaction compact_blocks(const ROperators& r, const std::vector<unsigned char>& keep_flat);

}  // namespace t3
