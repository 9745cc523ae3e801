// Effective-Hamiltonian plan builder; see heff_plan.hpp.
#include "heff_plan.hpp"
#include "host_threads.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace t3 {

namespace {

// Prefactors below this are exact zeros computed with rounding.
constexpr double kZeroCoupling = 1e-14;
const int kOrders[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

struct BlockIdx {
    QN qn[STEPSPECS_MSITES];
    int id[STEPSPECS_MSITES][3];
    const int* irr[STEPSPECS_MSITES][3];
    int dim[3];
};

// Sector indices, irreps and leg dimensions of one block
// (src/Heff.c:695-739).
void fill_block(const HeffInput& in, int isdmrg, int sb, BlockIdx* b) {
    const SiteTensor& t = *in.site;
    for (int i = 0; i < t.nrsites; ++i) {
        QN q = t.qnumbers[(size_t)t.nrsites * sb + i];
        b->qn[i] = q;
        const Symsecs* const* ss = in.symarr[i];
        b->id[i][0] = (int)(q % ss[0]->nrSecs);
        q /= ss[0]->nrSecs;
        b->id[i][1] = (int)(q % ss[1]->nrSecs);
        b->id[i][2] = (int)(q / ss[1]->nrSecs);
        for (int k = 0; k < 3; ++k) b->irr[i][k] = ss[k]->irreps[b->id[i][k]].data();
    }
    for (int i = 0; i < (isdmrg ? 2 : 3); ++i) {
        const int site = in.rops_on_site[i];
        const ROperators& op = *in.ops[i];
        if (!op.P_operator) {
            const int loose = isdmrg ? 2 * (i == 1) : i;
            b->dim[i] = in.symarr[site][loose]->dims[b->id[site][loose]];
        } else {
            const int loose = 2 * !op.is_left;
            b->dim[i] = in.symarr[site][loose]->dims[b->id[site][loose]] *
                        in.symarr[site][1]->dims[b->id[site][1]];
        }
    }
    if (isdmrg) b->dim[2] = 1;
}

int qn_cmp3(const QN* a, const QN* b) {
    for (int k = 2; k >= 0; --k)
        if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
    return 0;
}

// Operator block lookup (src/Heff.c:760-796).
int find_op_block(const HeffInput& in, int isdmrg, int i, const BlockIdx& nb, const BlockIdx& ob,
                  int idMPO) {
    const ROperators& op = *in.ops[i];
    const int site = in.rops_on_site[i];
    const int innerid = op.P_operator ? 2 * op.is_left : (isdmrg ? 2 * !op.is_left : i);
    const int dinner = in.symarr[site][innerid]->nrSecs;
    const QN qninner = (QN)nb.id[site][innerid] + (QN)ob.id[site][innerid] * dinner +
                       (QN)idMPO * dinner * dinner;
    const QN* arr = op.qnumbers_hss(idMPO);
    const int n = op.nblocks_hss(idMPO);
    if (!op.P_operator) {
        const QN* p = std::lower_bound(arr, arr + n, qninner);
        return (p != arr + n && *p == qninner) ? (int)(p - arr) : -1;
    }
    const QN key[3] = {nb.qn[site], ob.qn[site], qninner};
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        const int c = qn_cmp3(key, arr + 3 * (size_t)mid);
        if (c == 0) return mid;
        if (c < 0) hi = mid - 1;
        else lo = mid + 1;
    }
    return -1;
}

double calc_prefactor(const HeffInput& in, int isdmrg, int posB, const BlockIdx& nb,
                      const BlockIdx& ob, const int idMPO[3]) {
    // src/Heff.c:507-524
    double p = 1;
    for (int i = 0; i < (isdmrg ? 2 : 3); ++i) {
        const ROperators& op = *in.ops[i];
        if (!op.P_operator) continue;
        const int s = in.rops_on_site[i];
        const int* const irr[2][3] = {{nb.irr[s][0], nb.irr[s][1], nb.irr[s][2]},
                                      {ob.irr[s][0], ob.irr[s][1], ob.irr[s][2]}};
        p *= prefactor_add_P_operator(irr, op.is_left, in.sym);
    }
    const int* const irr[2][3] = {{nb.irr[posB][0], nb.irr[posB][1], nb.irr[posB][2]},
                                  {ob.irr[posB][0], ob.irr[posB][1], ob.irr[posB][2]}};
    const int* mpo[3] = {in.mpo->irreps[idMPO[0]].data(), in.mpo->irreps[idMPO[1]].data(),
                         isdmrg ? in.mpo->irreps[0].data() : in.mpo->irreps[idMPO[2]].data()};
    p *= prefactor_combine_MPOs(irr, mpo, in.sym, isdmrg, 2 * !in.ops[1]->P_operator);
    return p;
}

void mpo_ids_of_group(const HeffInput& in, int isdmrg, int group, int idMPO[3]) {
    const Instruction& ins = in.iset->instr[in.iset->MPOc_beg[group]];
    idMPO[0] = in.ops[0]->hss_of_ops[ins.instr[0]];
    idMPO[1] = in.ops[1]->hss_of_ops[ins.instr[1]];
    idMPO[2] = isdmrg ? 0 : in.ops[2]->hss_of_ops[ins.instr[2]];
    if (isdmrg && idMPO[0] != in.hermitian[idMPO[1]])
        throw std::runtime_error("Heff: inconsistent MPO sectors in DMRG merge");
}
}  // namespace

int heff_best_order(const int dold[3], const int dnew[3], int isdmrg) {
    int best = 0;
    long long bestops = 0;
    for (int i = 0; i < (isdmrg ? 2 : 6); ++i) {
        long long ops = 0;
        long long w[3] = {dold[0], dold[1], dold[2]};
        for (int j = 0; j < (isdmrg ? 2 : 3); ++j) {
            const int o = kOrders[i][j];
            ops += w[0] * w[1] * w[2] * dnew[o];
            w[o] = dnew[o];
        }
        if (i == 0 || bestops > ops) {
            best = i;
            bestops = ops;
        }
    }
    return best;
}

double heff_term_flops(const int dold[3], const int dnew[3], const int*, int order, int isdmrg) {
    double ops = 0;
    double w[3] = {(double)dold[0], (double)dold[1], (double)dold[2]};
    for (int j = 0; j < (isdmrg ? 2 : 3); ++j) {
        const int o = kOrders[order][j];
        ops += w[0] * w[1] * w[2] * dnew[o];
        w[o] = dnew[o];
    }
    return 2.0 * ops;
}

HeffPlan build_heff_plan(const HeffInput& in) {
    const SiteTensor& t = *in.site;
    HeffPlan P;
    P.iset = in.iset;
    const int ns = t.nrsites;
    P.isdmrg = (ns == 1 && in.site_is_phys[0]) ||
               (ns == 2 && in.site_is_phys[0] && in.site_is_phys[1]);
    const int isdmrg = P.isdmrg;
    P.posB = 0;
    for (int i = 0; i < ns; ++i)
        if (!in.site_is_phys[i]) P.posB = i;
    if (isdmrg) P.posB = ns - 1;
    const int posB = P.posB;
    // make_map (src/Heff.c:668-682)
    const bool swap01 = in.rops_on_site[1] != posB && in.rops_on_site[0] == posB;
    P.map[0] = swap01 ? 1 : 0;
    P.map[1] = swap01 ? 0 : 1;
    P.map[2] = 2;
    P.nrblocks = t.nrblocks;
    P.beginblock = t.beginblock;
    const int H = in.mpo->nrSecs;
    const int nops = isdmrg ? 2 : 3;

    // Unique qnB (branching-site or last-site quantum numbers), sorted.
    std::vector<QN> qnB(t.nrblocks);
    for (int i = 0; i < t.nrblocks; ++i) qnB[i] = t.qnumbers[(size_t)i * ns + posB];
    std::sort(qnB.begin(), qnB.end());
    qnB.erase(std::unique(qnB.begin(), qnB.end()), qnB.end());
    const int nqnB = (int)qnB.size();
    P.nr_qnB = nqnB;
    // Blocks per qnB (src/Heff.c:1138-1165).
    std::vector<std::vector<int>> sb_with(nqnB);
    for (int j = 0; j < t.nrblocks; ++j) {
        const QN q = t.qnumbers[(size_t)j * ns + posB];
        const int id = (int)(std::lower_bound(qnB.begin(), qnB.end(), q) - qnB.begin());
        sb_with[id].push_back(j);
    }

    // For every new qnB: the old qnBs and their raw MPO combinations.
    std::vector<std::vector<QN>> toqnB(nqnB);
    std::vector<std::vector<std::vector<int>>> combos(nqnB);
    if (isdmrg) {
        // src/Heff.c:362-442
        const int opid = in.ops[1]->P_operator;
        const ROperators& op = *in.ops[opid];
        if (!op.P_operator) throw std::runtime_error("DMRG Heff needs a physical operator");
        const int N = op.begin_blocks_of_hss[op.nrhss];
        for (int i = 0; i < nqnB; ++i) {
            int chss = 0;
            for (int j = 0; j < N; ++j) {
                while (op.begin_blocks_of_hss[chss + 1] <= j) ++chss;
                if (op.qnumbers[3 * (size_t)j] != qnB[i]) continue;
                const QN oq = op.qnumbers[3 * (size_t)j + 1];
                size_t k;
                for (k = 0; k < toqnB[i].size(); ++k)
                    if (toqnB[i][k] == oq) break;
                if (k == toqnB[i].size()) {
                    toqnB[i].push_back(oq);
                    combos[i].push_back({});
                }
                const int h1 = opid == 1 ? in.hermitian[chss] : chss;
                const int h2 = opid != 1 ? in.hermitian[chss] : chss;
                combos[i][k].push_back(h1 + h2 * H);
            }
        }
    } else {
        // Per operator: for every (bra, ket) sector pair of its branching
        // bond, the MPO sectors present (src/Heff.c:118-174).
        int idim[3];
        std::vector<std::vector<std::vector<int>>> qa[3];
        for (int k = 0; k < 3; ++k) {
            idim[k] = in.symarr[posB][k]->nrSecs;
            const ROperators& op = *in.ops[k];
            const int cpl = op.couplings();
            qa[k].assign(idim[k], std::vector<std::vector<int>>(idim[k]));
            QN prev = -1;
            int chss = 0;
            for (int i = 0; i < op.begin_blocks_of_hss[op.nrhss]; ++i) {
                const QN cq = op.qnumbers[(size_t)cpl * i + cpl - 1];
                while (i >= op.begin_blocks_of_hss[chss + 1]) ++chss;
                if (prev == cq) continue;
                const int bra = (int)(cq % idim[k]);
                const int ket = (int)((cq / idim[k]) % idim[k]);
                qa[k][bra][ket].push_back(chss);
                prev = cq;
            }
        }
        // Enumerate old qnBs (src/Heff.c:181-360); old index pairs (i0, i1)
        // in the order of find_next_old, third index from qnB itself.
        const QN big = (QN)idim[0] * idim[1];
        for (int i = 0; i < nqnB; ++i) {
            int idn[3];
            QN q = qnB[i];
            idn[0] = (int)(q % idim[0]);
            q /= idim[0];
            idn[1] = (int)(q % idim[1]);
            idn[2] = (int)(q / idim[1]);
            const auto& r0 = qa[0][idn[0]];
            const auto& r1 = qa[1][idn[1]];
            const auto& r2 = qa[2][idn[2]];
            for (int i1 = 0; i1 < idim[1]; ++i1) {
                if (r1[i1].empty()) continue;
                for (int i0 = 0; i0 < idim[0]; ++i0) {
                    if (r0[i0].empty()) continue;
                    const QN key = (QN)i0 + (QN)i1 * idim[0];
                    for (int loc = 0; loc < nqnB; ++loc) {
                        if (qnB[loc] % big != key) continue;
                        const int i2 = (int)(qnB[loc] / big);
                        std::vector<int> cmb;
                        if (!r2[i2].empty()) {
                            for (int m0 : r0[i0])
                                for (int m1 : r1[i1])
                                    for (int m2 : r2[i2]) {
                                        const int h2 = in.hermitian[m2];
                                        if (in.singlet[(size_t)m0 + (size_t)m1 * H + (size_t)h2 * H * H])
                                            cmb.push_back(m0 + m1 * H + m2 * H * H);
                                    }
                        }
                        if (!cmb.empty()) {
                            toqnB[i].push_back(qnB[loc]);
                            combos[i].push_back(std::move(cmb));
                        }
                    }
                }
            }
        }
    }
    // MPO combination -> merge group (src/Heff.c:1110-1136).
    const auto& MPOc = in.iset->MPOc;
    for (int i = 0; i < nqnB; ++i)
        for (auto& cl : combos[i]) {
            std::vector<int> kept;
            for (int v : cl) {
                auto it = std::find(MPOc.begin(), MPOc.end(), v);
                if (it != MPOc.end()) kept.push_back((int)(it - MPOc.begin()));
            }
            cl = std::move(kept);
        }

    // Block dims.
    P.dims.assign(t.nrblocks, {0, 0, 0});
    std::vector<BlockIdx> bidx(t.nrblocks);
    for (int sb = 0; sb < t.nrblocks; ++sb) {
        fill_block(in, isdmrg, sb, &bidx[sb]);
        for (int k = 0; k < 3; ++k) P.dims[sb][k] = bidx[sb].dim[k];
    }

    // First-run table (src/Heff.c:815-914, 1002-1047).  Every qnB owns its
    // new blocks' entries of bynew, so host threads take qnBs from a shared
    // counter; the pairs are joined in new-block order afterwards.
    std::vector<std::vector<PlanPair>> bynew(t.nrblocks);
    run_blocks_threaded(nqnB, [&](int nq) {
        for (int newsb : sb_with[nq]) {
            const BlockIdx& nb = bidx[newsb];
            for (size_t oq = 0; oq < toqnB[nq].size(); ++oq) {
                const int oid = (int)(std::lower_bound(qnB.begin(), qnB.end(), toqnB[nq][oq]) - qnB.begin());
                const auto& mpos = combos[nq][oq];
                for (int oldsb : sb_with[oid]) {
                    const BlockIdx& ob = bidx[oldsb];
                    PlanPair pp;
                    pp.newsb = newsb;
                    pp.oldsb = oldsb;
                    pp.bestorder = heff_best_order(ob.dim, nb.dim, isdmrg);
                    for (int g : mpos) {
                        const int beg = in.iset->MPOc_beg[g], end = in.iset->MPOc_beg[g + 1];
                        if (end == beg) continue;
                        int idMPO[3];
                        mpo_ids_of_group(in, isdmrg, g, idMPO);
                        const double pref = calc_prefactor(in, isdmrg, posB, nb, ob, idMPO);
                        // The reference drops |pref| < 1e-20 (src/Heff.c:827), which keeps
                        // or drops a coupling that is exactly zero depending on the rounding
                        // of its 9j evaluation (values ~1e-18).  Nonzero couplings at the
                        // spins that occur are many orders larger, so everything below
                        // kZeroCoupling is an exact zero and its GEMMs are skipped.
                        if (std::fabs(pref) < kZeroCoupling) continue;
                        PlanEntry e;
                        e.mpo_group = g;
                        e.prefactor = pref;
                        e.sbops[2] = -1;
                        bool ok = true;
                        for (int k = 0; k < nops; ++k) {
                            e.sbops[k] = find_op_block(in, isdmrg, k, nb, ob, idMPO[k]);
                            if (e.sbops[k] < 0 && k < 2) ok = false;
                        }
                        if (!ok) continue;
                        if (!isdmrg && e.sbops[2] < 0)
                            throw std::runtime_error("Heff: operator block not found");
                        pp.entries.push_back(e);
                    }
                    if (!pp.entries.empty()) bynew[newsb].push_back(std::move(pp));
                }
            }
        }
    }, (double)t.nrblocks * t.nrblocks, 1e6);  // pairs considered, at most
    P.pair_begin.assign(t.nrblocks + 1, 0);
    for (int sb = 0; sb < t.nrblocks; ++sb) {
        P.pair_begin[sb + 1] = P.pair_begin[sb] + (int)bynew[sb].size();
        for (auto& pp : bynew[sb]) P.pairs.push_back(std::move(pp));
    }
    // Algorithmic flops: executed GEMMs only (empty operator blocks skipped,
    // src/Heff.c:798-813, 926-933).
    for (const auto& pp : P.pairs) {
        const int* dn = P.dims[pp.newsb].data();
        const int* dd = P.dims[pp.oldsb].data();
        const double f = heff_term_flops(dd, dn, P.map, pp.bestorder, isdmrg);
        for (const auto& e : pp.entries) {
            for (int i = in.iset->MPOc_beg[e.mpo_group]; i < in.iset->MPOc_beg[e.mpo_group + 1]; ++i) {
                const Instruction& ins = in.iset->instr[i];
                bool ok = true;
                for (int k = 0; k < nops; ++k)
                    if (in.ops[k]->block_size(ins.instr[k], e.sbops[k]) == 0) ok = false;
                if (!ok) continue;
                P.flops += f;
                ++P.nterms;
            }
        }
    }
    return P;
}

std::vector<DiagEntry> build_heff_diag_plan(const HeffInput& in, const HeffPlan& P) {
    // For each block, the MPO groups of its diagonal pair (src/Heff.c:1268-1312).
    std::vector<DiagEntry> out;
    const int isdmrg = P.isdmrg;
    const int nops = isdmrg ? 2 : 3;
    for (int sb = 0; sb < P.nrblocks; ++sb) {
        BlockIdx b;
        fill_block(in, isdmrg, sb, &b);
        // The diagonal pair's groups: take them from the plan when the pair
        // survived, otherwise recompute (all-zero prefactors are dropped from
        // the plan but kept here).
        for (int pi = P.pair_begin[sb]; pi < P.pair_begin[sb + 1]; ++pi) {
            const PlanPair& pp = P.pairs[pi];
            if (pp.oldsb != sb) continue;
            for (const auto& e : pp.entries) {
                DiagEntry d;
                d.sb = sb;
                d.mpo_group = e.mpo_group;
                for (int k = 0; k < 3; ++k) d.sbops[k] = k < nops ? e.sbops[k] : -1;
                d.prefactor = e.prefactor;
                out.push_back(d);
            }
        }
    }
    return out;
}

}  // namespace t3
