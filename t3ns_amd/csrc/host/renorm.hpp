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

// Drops the blocks flagged empty: returns the compacted per-operator block
// offsets and, for every kept block, (old offset, new offset, size).
struct Compaction {
    std::vector<std::vector<int64_t>> beginblock;
    std::vector<int64_t> op_offset;
    std::vector<int64_t> src, dst, n;
};
Compaction compact_blocks(const ROperators& r, const std::vector<unsigned char>& keep_flat);

}  // namespace t3
