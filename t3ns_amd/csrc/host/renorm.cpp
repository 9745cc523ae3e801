// Append plan (src/rOperators_pUpdate.c:393-673, src/rOperators_init.c:185-265).
#include "renorm.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>
#include <string>

#include "host_threads.hpp"

namespace t3 {

namespace {

void indexize3(int* id, QN qn, const Symsecs* ss) {
    const QN n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
    id[0] = (int)(qn % n0);
    id[1] = (int)((qn / n0) % n1);
    id[2] = (int)(qn / (n0 * n1));
}

QN qntypize3(const int* id, const Symsecs* ss) {
    return (QN)id[0] + (QN)id[1] * ss[0].nrSecs + (QN)id[2] * ss[0].nrSecs * ss[1].nrSecs;
}

int find_qn(const QN* arr, int n, QN v) {
    const QN* p = std::lower_bound(arr, arr + n, v);
    return (p != arr + n && *p == v) ? (int)(p - arr) : -1;
}

}  // namespace

AppendPlan plan_append(const AppendInput& in) {
    const ROperators& old = *in.old;
    if (old.P_operator) throw std::invalid_argument("append: old operators are already physical");
    AppendPlan P;
    ROperators& nr = P.nr;
    nr.bond = in.new_ss[2].bond;
    nr.is_left = in.is_left;
    nr.P_operator = 1;
    nr.nrhss = in.nrhss;
    nr.begin_blocks_of_hss = in.begin_blocks_of_hss;
    nr.qnumbers = in.qnumbers;
    // new operators: one per distinct instr[2] (src/rOperators_init.c:185-226)
    int nops = 0;
    for (const auto& i : in.instr) nops = std::max(nops, i.instr[2] + 1);
    nr.nrops = nops;
    nr.hss_of_ops.assign(in.hss_of_new.begin(), in.hss_of_new.begin() + nops);
    nr.beginblock.resize(nops);
    for (int k = 0; k < nops; ++k) nr.beginblock[k] = in.hss_bb[nr.hss_of_ops[k]];
    nr.finalize_offsets();

    // unique operators: runs of equal (old op, site op, MPO sector of new)
    // in the sorted instructions (src/instructions.c:224-252)
    const int ni = (int)in.instr.size();
    std::vector<int> uniq(ni);
    struct U {
        int po, so, hssn, hsso, hsss;
    };
    std::vector<U> us;
    for (int i = 0; i < ni; ++i) {
        const Instruction& a = in.instr[i];
        const bool fresh = i == 0 || a.instr[0] != in.instr[i - 1].instr[0] ||
                           a.instr[1] != in.instr[i - 1].instr[1] ||
                           in.hss_of_new[a.instr[2]] != in.hss_of_new[in.instr[i - 1].instr[2]];
        if (fresh) {
            U u;
            u.po = a.instr[0];
            u.so = a.instr[1];
            u.hssn = in.hss_of_new[a.instr[2]];
            u.hsso = old.hss_of_ops[u.po];
            u.hsss = in.siteop_hss[u.so];
            us.push_back(u);
        }
        uniq[i] = (int)us.size() - 1;
    }

    // pAppend_block: for every new block, the old block and factor of each
    // unique operator that writes it (src/rOperators_pUpdate.c:537-640)
    struct Src {
        int64_t off = -1;
        double factor = 0;
    };
    // src_of[u][local block]
    std::vector<std::vector<Src>> src_of(us.size());
    for (size_t u = 0; u < us.size(); ++u) src_of[u].assign(nr.nblocks_hss(us[u].hssn), Src{});
    const int pbond = 2 * !in.is_left;
    const int nblocks = nr.begin_blocks_of_hss[nr.nrhss];
    for (int usb = 0; usb < nblocks; ++usb) {
        int ids[3][3];
        indexize3(ids[0], in.qnumbers[3 * usb + 0], in.site_ss);
        indexize3(ids[1], in.qnumbers[3 * usb + 1], in.site_ss);
        indexize3(ids[2], in.qnumbers[3 * usb + 2], in.new_ss);
        const int hssn = ids[2][2];
        ids[2][pbond == 0 ? 2 : 0] = hssn;
        const int ublock = usb - nr.begin_blocks_of_hss[hssn];
        for (int t = in.tprod_begin[hssn]; t < in.tprod_begin[hssn + 1]; ++t) {
            const int hsss = in.tprods[2 * t], hsso = in.tprods[2 * t + 1];
            ids[2][pbond] = hsso;
            ids[2][1] = hsss;
            const int id_old[3] = {ids[0][pbond], ids[1][pbond], ids[2][pbond]};
            const QN oqn = qntypize3(id_old, in.old_ss);
            const int* irr[3][3];
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 3; ++j) irr[i][j] = in.site_ss[j].irreps[ids[i][j]].data();
            for (int j = 0; j < 3; ++j) irr[2][j] = in.mpo.irreps[ids[2][j]].data();
            const double pref = prefactor_pAppend(irr, in.is_left, in.sym);
            const int oblock = find_qn(old.qnumbers_hss(hsso), old.nblocks_hss(hsso), oqn);
            if (oblock == -1 || std::fabs(pref) < 1e-20) continue;
            for (size_t u = 0; u < us.size(); ++u) {
                const U& uu = us[u];
                if (!(uu.hsso == hsso && uu.hsss == hsss && uu.hssn == hssn)) continue;
                if (old.block_size(uu.po, oblock) == 0) continue;
                const double el = in.siteop_el[((size_t)uu.so * in.physdim + ids[0][1]) * in.physdim +
                                               ids[1][1]];
                Src& s = src_of[u][ublock];
                s.off = old.op_offset[uu.po] + old.beginblock[uu.po][oblock];
                s.factor = pref * el;
            }
        }
    }

    // sum_unique_rOperators: new op k, block b = sum over instructions with
    // target k of pref * unique(b), in instruction order
    // (src/rOperators_init.c:228-258)
    std::vector<std::vector<int>> instr_of(nops);
    for (int i = 0; i < ni; ++i) instr_of[in.instr[i].instr[2]].push_back(i);
    for (int k = 0; k < nops; ++k) {
        const int nb = nr.nblocks_op(k);
        for (int b = 0; b < nb; ++b) {
            AppendGroup g;
            g.dst = nr.op_offset[k] + nr.beginblock[k][b];
            g.n = nr.block_size(k, b);
            g.op = k;
            g.block = b;
            g.task_begin = (int)P.tasks.size();
            for (int i : instr_of[k]) {
                const Src& s = src_of[uniq[i]][b];
                if (s.off < 0) continue;
                P.tasks.push_back(AppendTask{s.off, s.factor, in.instr[i].pref});
            }
            g.task_end = (int)P.tasks.size();
            if (g.task_end > g.task_begin && g.n > 0) P.groups.push_back(g);
            else P.tasks.resize(g.task_begin);
        }
    }
    return P;
}

Compaction compact_blocks(const ROperators& r, const std::vector<unsigned char>& keep) {
    Compaction c;
    c.beginblock.resize(r.nrops);
    c.op_offset.assign(r.nrops + 1, 0);
    size_t flat = 0;
    for (int k = 0; k < r.nrops; ++k) {
        const int nb = r.nblocks_op(k);
        auto& bb = c.beginblock[k];
        bb.assign(nb + 1, 0);
        for (int b = 0; b < nb; ++b, ++flat) {
            const int64_t n = keep[flat] ? r.block_size(k, b) : 0;
            bb[b + 1] = bb[b] + n;
            if (n) {
                c.src.push_back(r.op_offset[k] + r.beginblock[k][b]);
                c.dst.push_back(c.op_offset[k] + bb[b]);
                c.n.push_back(n);
            }
        }
        c.op_offset[k + 1] = c.op_offset[k] + bb[nb];
    }
    return c;
}

}  // namespace t3

namespace t3 {

namespace {
int search_sector(const int* irrep, const Symsecs& ss, int nrsyms) {
    for (int i = 0; i < ss.nrSecs; ++i) {
        int k = 0;
        for (; k < nrsyms; ++k)
            if (ss.irreps[i][k] != irrep[k]) break;
        if (k == nrsyms) return i;
    }
    return -1;
}
}  // namespace

PUpdatePlan plan_pupdate(const PUpdateInput& in) {
    const ROperators& po = *in.pops;
    if (!po.P_operator) throw std::invalid_argument("pUpdate: operators are not physical");
    const SiteTensor& T = *in.T;
    const int il = in.is_left;
    const int ns = in.sym.nrSyms;
    PUpdatePlan P;
    ROperators& nr = P.nr;
    nr.bond = po.bond;
    nr.is_left = il;
    nr.P_operator = 0;
    nr.nrhss = in.nrhss;
    nr.begin_blocks_of_hss = in.begin_blocks_of_hss;
    nr.qnumbers = in.qnumbers;
    nr.nrops = po.nrops;
    nr.hss_of_ops = po.hss_of_ops;
    nr.beginblock.resize(nr.nrops);
    for (int k = 0; k < nr.nrops; ++k) nr.beginblock[k] = in.hss_bb[nr.hss_of_ops[k]];
    nr.finalize_offsets();

    // uss: (new bond, new bond, MPO); oss[0] = oss[1]: site bonds with the
    // updated bond internal; oss[2] = (internal, internal, MPO)
    // (src/rOperators_pUpdate.c:140-176)
    const int ub = il ? 2 : 0;
    Symsecs uss[3] = {in.site_ss[ub], in.site_ss[ub], in.mpo};
    Symsecs oss0[3] = {in.site_ss[0], in.site_ss[1], in.site_ss[2]};
    oss0[ub] = in.iss;
    Symsecs oss2[3] = {in.iss, in.iss, in.mpo};

    // make_usb_to_osb: range of old blocks per new block
    const int N = nr.begin_blocks_of_hss[nr.nrhss];
    const int NN = po.begin_blocks_of_hss[po.nrhss];
    std::vector<int> usb_to_osb(N + 1, NN);
    usb_to_osb[0] = 0;
    {
        int osb = 0;
        for (int usb = 0; usb < N; ++usb) {
            int id[3], nid[3];
            indexize3(id, nr.qnumbers[usb], uss);
            for (int k = 0; k < 2; ++k) nid[k] = search_sector(uss[k].irreps[id[k]].data(), oss2[k], ns);
            nid[2] = id[2];
            const QN qn = qntypize3(nid, oss2);
            for (; osb < NN; ++osb)
                if (qn < po.qnumbers[3 * osb + 2]) {
                    usb_to_osb[usb + 1] = osb;
                    break;
                }
        }
    }
    std::vector<QN> Tqn(T.qnumbers.begin(), T.qnumbers.end());
    for (int usb = 0; usb < N; ++usb) {
        int newhss = 0;
        while (usb >= nr.begin_blocks_of_hss[newhss + 1]) ++newhss;
        const int usb2 = usb - nr.begin_blocks_of_hss[newhss];
        // per old block: site blocks, dims and prefactor (select_site_blocks)
        struct Aide {
            int osb2;
            int64_t th, t;
            int M0, M1, N0, N1;
            double pref;
        };
        std::vector<Aide> aides;
        for (int osb = usb_to_osb[usb]; osb < usb_to_osb[usb + 1]; ++osb) {
            const QN* cqn = &po.qnumbers[3 * osb];
            int ids[3][3];
            indexize3(ids[0], cqn[0], oss0);
            indexize3(ids[1], cqn[1], oss0);
            indexize3(ids[2], cqn[2], oss2);
            int Tids[2][3];
            bool ok = true;
            for (int k = 0; k < 3; ++k) {
                Tids[0][k] = search_sector(oss0[k].irreps[ids[1][k]].data(), in.site_ss[k], ns);
                Tids[1][k] = search_sector(oss0[k].irreps[ids[0][k]].data(), in.site_ss[k], ns);
                ok = ok && Tids[0][k] >= 0 && Tids[1][k] >= 0;
            }
            if (!ok) continue;
            const QN tq = qntypize3(Tids[0], in.site_ss), thq = qntypize3(Tids[1], in.site_ss);
            const int tsb = find_qn(Tqn.data(), T.nrblocks, tq);
            const int thsb = find_qn(Tqn.data(), T.nrblocks, thq);
            if (tsb < 0 || thsb < 0) continue;
            if (T.block_size(tsb) == 0 || T.block_size(thsb) == 0) continue;
            Aide a;
            a.osb2 = osb - po.begin_blocks_of_hss[newhss];
            a.th = T.beginblock[thsb];
            a.t = T.beginblock[tsb];
            a.N0 = in.site_ss[2 * il].dims[Tids[1][2 * il]];
            a.N1 = in.site_ss[2 * il].dims[Tids[0][2 * il]];
            a.M0 = in.site_ss[2 * !il].dims[Tids[1][2 * !il]] * in.site_ss[1].dims[Tids[1][1]];
            a.M1 = in.site_ss[2 * !il].dims[Tids[0][2 * !il]] * in.site_ss[1].dims[Tids[0][1]];
            const int* irr[3][3];
            irr[0][0] = oss0[0].irreps[ids[0][0]].data();
            irr[0][1] = oss0[1].irreps[ids[0][1]].data();
            irr[0][2] = oss0[2].irreps[ids[0][2]].data();
            irr[1][0] = oss0[0].irreps[ids[1][0]].data();
            irr[1][1] = oss0[1].irreps[ids[1][1]].data();
            irr[1][2] = oss0[2].irreps[ids[1][2]].data();
            for (int j = 0; j < 3; ++j) irr[2][j] = oss2[j].irreps[ids[2][j]].data();
            const int* adj[3] = {irr[0][0], irr[0][1], irr[0][2]};
            a.pref = prefactor_adjoint(adj, il ? '3' : '1', in.sym) *
                     prefactor_pUpdate(irr, il, in.sym);
            aides.push_back(a);
        }
        for (int k = 0; k < po.nrops; ++k) {
            if (po.hss_of_ops[k] != newhss) continue;
            PUpdateGroup g;
            g.op = k;
            g.block = usb2;
            g.dst = nr.op_offset[k] + nr.beginblock[k][usb2];
            g.n = nr.block_size(k, usb2);
            g.task_begin = (int)P.tasks.size();
            if (g.n > 0) {
                for (const Aide& a : aides) {
                    if (po.block_size(k, a.osb2) == 0) continue;
                    if ((int64_t)a.N0 * a.N1 != g.n ||
                        (int64_t)a.M0 * a.M1 != po.block_size(k, a.osb2))
                        throw std::runtime_error("pUpdate: block size mismatch");
                    PUpdateTask t;
                    t.dst = g.dst;
                    t.src = po.op_offset[k] + po.beginblock[k][a.osb2];
                    t.th = a.th;
                    t.t = a.t;
                    t.M0 = a.M0;
                    t.M1 = a.M1;
                    t.N0 = a.N0;
                    t.N1 = a.N1;
                    t.pref = a.pref;
                    P.tasks.push_back(t);
                }
            }
            g.task_end = (int)P.tasks.size();
            if (g.task_end > g.task_begin) P.groups.push_back(g);
        }
    }
    return P;
}

}  // namespace t3

namespace t3 {

// update_rOperators_branching (src/rOperators_bUpdate.c:229-248): the
// enumeration of update_unique_ops_T3NS (637-677), update_newblock_w_MPO_set
// (746-785), next_sb_tens (680-718), next_sb_sec_op (720-744),
// find_block_adj (787-806) and update_selected_blocks (1043-1072), then
// sum_unique_rOperators (src/rOperators_init.c:228-265).
// The branching update is one small GEMM chain per (new block, tensor
// block, adjoint block, instruction); with many tiny symmetry sectors (e.g.
// 36 orbitals under D2h) their number reaches 1e8-1e9, whose task list
// alone outgrows host memory.  Past this count the planner stops with an
// error instead.
constexpr int64_t kMaxBUpdateTasks = 100000000;



BUpdatePlan plan_bupdate(const BUpdateInput& in) {
    const int uc = in.update_case;
    if (uc < 0 || uc > 2) throw std::invalid_argument("bUpdate: bad update case");
    const ROperators& O0 = *in.ops[0];
    const ROperators& O1 = *in.ops[1];
    if (O0.P_operator || O1.P_operator) throw std::invalid_argument("bUpdate: physical operators");
    const SiteTensor& T = *in.T;
    const int H = in.mpo.nrSecs;
    BUpdatePlan P;
    // bond of each role: ops[0], ops[1], new (initialize_indexhelper, 518-520)
    const int id_ops[3] = {uc == 0 ? 1 : 0, uc == 2 ? 1 : 2, uc};
    P.x_leg = id_ops[0];
    P.y_leg = id_ops[1];
    P.n_leg = uc;
    const int looptype = uc == 0;
    const int nsec[3] = {in.site_ss[0].nrSecs, in.site_ss[1].nrSecs, in.site_ss[2].nrSecs};

    // unique operators: runs of equal (op0, op1, MPO sector of the new op)
    const int ni = (int)in.instr.size();
    std::vector<int> run_first, run_of(ni);
    for (int i = 0; i < ni; ++i) {
        const Instruction& a = in.instr[i];
        const bool fresh = i == 0 || a.instr[0] != in.instr[i - 1].instr[0] ||
                           a.instr[1] != in.instr[i - 1].instr[1] ||
                           in.hss_of_new[a.instr[2]] != in.hss_of_new[in.instr[i - 1].instr[2]];
        if (fresh) run_first.push_back(i);
        run_of[i] = (int)run_first.size() - 1;
    }
    const int nu = (int)run_first.size();
    ROperators& U = P.uniq;
    U.bond = -1;
    U.is_left = uc == 2;
    U.P_operator = 0;
    U.nrhss = in.nrhss;
    U.begin_blocks_of_hss = in.begin_blocks_of_hss;
    U.qnumbers = in.qnumbers;
    U.nrops = nu;
    U.hss_of_ops.resize(nu);
    U.beginblock.resize(nu);
    for (int u = 0; u < nu; ++u) {
        U.hss_of_ops[u] = in.hss_of_new[in.instr[run_first[u]].instr[2]];
        U.beginblock[u] = in.hss_bb[U.hss_of_ops[u]];
    }
    U.finalize_offsets();

    // instruction helper: MPO combination -> (first instruction, unique id)
    // (init_instrhelper, 423-493)
    std::vector<std::pair<QN, std::vector<std::array<int, 2>>>> combos;
    for (int u = 0; u < nu; ++u) {
        const int* ins = in.instr[run_first[u]].instr;
        if (ins[0] < 0 || ins[0] >= O0.nrops || ins[1] < 0 || ins[1] >= O1.nrops)
            throw std::invalid_argument("bUpdate: instruction out of range");
        const QN key = (QN)O0.hss_of_ops[ins[0]] + (QN)O1.hss_of_ops[ins[1]] * H +
                       (QN)in.hss_of_new[ins[2]] * H * H;
        auto it = std::find_if(combos.begin(), combos.end(), [&](const auto& c) { return c.first == key; });
        if (it == combos.end()) {
            combos.push_back({key, {}});
            it = combos.end() - 1;
        }
        it->second.push_back({run_first[u], u});
    }
    std::sort(combos.begin(), combos.end(), [](const auto& a, const auto& b) { return a.first < b.first; });

    // tensor blocks grouped by the two kets the loop fixes (make_qntens, 304-383)
    const QN divide = looptype ? (QN)nsec[0] * nsec[1] : (QN)nsec[0];
    std::vector<std::pair<QN, std::vector<int>>> tgroups;
    for (int b = 0; b < T.nrblocks; ++b) {
        const QN key = looptype ? T.qnumbers[b] % divide : T.qnumbers[b] / divide;
        auto it = std::find_if(tgroups.begin(), tgroups.end(), [&](const auto& g) { return g.first == key; });
        if (it == tgroups.end()) {
            tgroups.push_back({key, {}});
            it = tgroups.end() - 1;
        }
        it->second.push_back(b);
    }
    std::sort(tgroups.begin(), tgroups.end(), [](const auto& a, const auto& b) { return a.first < b.first; });

    const ROperators* ops[2] = {&O0, &O1};
    auto ids_of = [&](QN qn, int bond, int* bra, int* ket, int* h) {
        const QN n = nsec[bond];
        *bra = (int)(qn % n);
        *ket = (int)((qn / n) % n);
        *h = (int)(qn / (n * n));
    };
    const int first = looptype ? 0 : 1, second = 1 - first;
    const int N = in.begin_blocks_of_hss[in.nrhss];
    // Contributions of every new block, grouped by unique operator u (in
    // loop order within a group): each new block fills only its own entry,
    // so host threads take new blocks from a shared counter; the
    // destinations (u, block) are then laid out in (u, block) order, which
    // makes the plan independent of the threads.
    std::vector<std::vector<BUpdateTask>> blk_tasks(N);
    std::vector<std::vector<int>> blk_uoff(N);  // [nu + 1] offsets per new block
    std::atomic<int64_t> ntasks{0};
    auto plan_block = [&](int new_sb) {
        std::vector<BUpdateTask> loc;
        std::vector<int> loc_u;
        int bra[3], ket[3], hss[3];
        ids_of(in.qnumbers[new_sb], uc, &bra[uc], &ket[uc], &hss[2]);
        const int hn = hss[2];
        const int nsb = new_sb - in.begin_blocks_of_hss[hn];
        for (int p = in.tprod_begin[hn]; p < in.tprod_begin[hn + 1]; ++p) {
            const int hss_ops[2] = {in.tprods[2 * p], in.tprods[2 * p + 1]};
            const ROperators& F = *ops[first];
            const ROperators& S = *ops[second];
            const int bf = id_ops[first], bs = id_ops[second];
            const int nbf = F.nblocks_hss(hss_ops[first]);
            const QN* qf = F.qnumbers_hss(hss_ops[first]);
            for (int jf = 0; jf < nbf; ++jf) {
                int hf;
                ids_of(qf[jf], bf, &bra[bf], &ket[bf], &hf);
                const QN match = looptype ? (QN)ket[0] + (QN)ket[1] * nsec[0]
                                          : (QN)ket[1] + (QN)ket[2] * nsec[1];
                const auto tg = std::lower_bound(tgroups.begin(), tgroups.end(), match,
                                                 [](const auto& g, QN v) { return g.first < v; });
                if (tg == tgroups.end() || tg->first != match) continue;
                const int nbs = S.nblocks_hss(hss_ops[second]);
                const QN* qs = S.qnumbers_hss(hss_ops[second]);
                // the second operator's blocks of one MPO sector are sorted by
                // (bra + ket n + hss n^2): the blocks with a given ket form
                // one contiguous run, found by bisection (scanned in order,
                // as a full scan would visit them)
                const QN nb2 = nsec[bs];
                bool runs_ok = nbs > 0;
                for (int js = 1; js < nbs && runs_ok; ++js) runs_ok = qs[js - 1] < qs[js];
                const QN h0 = nbs > 0 ? qs[0] / (nb2 * nb2) : 0;
                if (runs_ok) runs_ok = qs[nbs - 1] / (nb2 * nb2) == h0;
                for (int tb : tg->second) {
                    ket[bs] = looptype ? (int)(T.qnumbers[tb] / divide) : (int)(T.qnumbers[tb] % divide);
                    int js_lo = 0, js_hi = nbs;
                    if (runs_ok) {
                        const QN lo = (QN)ket[bs] * nb2 + h0 * nb2 * nb2;
                        js_lo = (int)(std::lower_bound(qs, qs + nbs, lo) - qs);
                        js_hi = (int)(std::lower_bound(qs + js_lo, qs + nbs, lo + nb2) - qs);
                    }
                    for (int js = js_lo; js < js_hi; ++js) {
                        int ks, hs;
                        ids_of(qs[js], bs, &bra[bs], &ks, &hs);
                        if (ks != ket[bs]) continue;
                        const QN aqn = (QN)bra[0] + (QN)bra[1] * nsec[0] + (QN)bra[2] * nsec[0] * nsec[1];
                        const int ab = find_qn(T.qnumbers.data(), T.nrblocks, aqn);
                        if (ab < 0) continue;
                        int sbop[2];
                        sbop[first] = jf;
                        sbop[second] = js;
                        // prefactor over (bra, ket, MPO) irreps of the three bonds
                        const int hb[3] = {hss_ops[0], hss_ops[1], hn};
                        int hsite[3];
                        for (int r = 0; r < 3; ++r) hsite[id_ops[r]] = hb[r];
                        const int* irr[3][3];
                        for (int i = 0; i < 3; ++i) {
                            irr[i][0] = in.site_ss[i].irreps[bra[i]].data();
                            irr[i][1] = in.site_ss[i].irreps[ket[i]].data();
                            irr[i][2] = in.mpo.irreps[hsite[i]].data();
                        }
                        const double pref = prefactor_bUpdate(irr, uc, in.sym);
                        // an exactly zero prefactor adds nothing
                        if (pref == 0.0) continue;
                        const QN key = (QN)hss_ops[0] + (QN)hss_ops[1] * H + (QN)hn * H * H;
                        const auto cb = std::lower_bound(combos.begin(), combos.end(), key,
                                                         [](const auto& c, QN v) { return c.first < v; });
                        if (cb == combos.end() || cb->first != key) continue;
                        const int64_t tsz = T.block_size(tb), asz = T.block_size(ab);
                        if (tsz == 0 || asz == 0) continue;
                        for (const auto& iu : cb->second) {
                            const int* ins = in.instr[iu[0]].instr;
                            const int u = iu[1];
                            if (O0.block_size(ins[0], sbop[0]) == 0 || O1.block_size(ins[1], sbop[1]) == 0 ||
                                U.block_size(u, nsb) == 0)
                                continue;
                            BUpdateTask t;
                            t.t = T.beginblock[tb];
                            t.a = T.beginblock[ab];
                            t.o1 = O0.op_offset[ins[0]] + O0.beginblock[ins[0]][sbop[0]];
                            t.o2 = O1.op_offset[ins[1]] + O1.beginblock[ins[1]][sbop[1]];
                            for (int i = 0; i < 3; ++i) {
                                t.d[i] = in.site_ss[i].dims[ket[i]];
                                t.e[i] = in.site_ss[i].dims[bra[i]];
                            }
                            t.pref = pref;
                            loc.push_back(t);
                            loc_u.push_back(u);
                        }
                    }
                }
            }
        }
        const int64_t added = (int64_t)loc.size();
        if (ntasks.fetch_add(added) + added > kMaxBUpdateTasks)
            throw std::runtime_error(
                "bUpdate: more than " + std::to_string(kMaxBUpdateTasks) +
                " block products (tiny symmetry sectors); the plan would not fit in host memory");
        // counting sort by u (stable: loop order kept within each u)
        std::vector<int>& off = blk_uoff[new_sb];
        off.assign(nu + 1, 0);
        for (int u : loc_u) ++off[u + 1];
        for (int u = 0; u < nu; ++u) off[u + 1] += off[u];
        std::vector<int> pos(off.begin(), off.end() - 1);
        std::vector<BUpdateTask>& out = blk_tasks[new_sb];
        out.resize(loc.size());
        for (size_t i = 0; i < loc.size(); ++i) out[pos[loc_u[i]]++] = loc[i];
    };
    run_blocks_threaded(N, plan_block, (double)N * T.nrblocks, 1e5);  // block combinations, roughly
    // destinations in (u, block) order, then the tasks copied into place
    struct Seg {
        int sb;
        int64_t from, to;
    };
    std::vector<Seg> segs;
    int64_t total = 0;
    for (int u = 0; u < nu; ++u) {
        const int h = U.hss_of_ops[u];
        for (int b = 0; b < U.nblocks_hss(h); ++b) {
            const int sb = in.begin_blocks_of_hss[h] + b;
            const std::vector<int>& off = blk_uoff[sb];
            if (off.empty() || off[u + 1] == off[u]) continue;
            const BUpdateTask& t0 = blk_tasks[sb][off[u]];
            BUpdateDest dd;
            dd.dst = U.op_offset[u] + U.beginblock[u][b];
            dd.M = t0.e[uc];
            dd.N = t0.d[uc];
            dd.task_begin = (int)total;
            total += off[u + 1] - off[u];
            dd.task_end = (int)total;
            P.dests.push_back(dd);
            segs.push_back(Seg{sb, off[u], off[u + 1]});
        }
    }
    if (total > INT32_MAX) throw std::runtime_error("bUpdate: more than 2^31 block products");
    P.tasks.resize(total);
    run_blocks_threaded((int)segs.size(), [&](int i) {
        const Seg& g = segs[i];
        std::copy(blk_tasks[g.sb].begin() + g.from, blk_tasks[g.sb].begin() + g.to,
                  P.tasks.begin() + P.dests[i].task_begin);
    }, (double)total, 1e6);

    // sum_unique_rOperators
    ROperators& nr = P.nr;
    nr.bond = -1;
    nr.is_left = U.is_left;
    nr.P_operator = 0;
    nr.nrhss = in.nrhss;
    nr.begin_blocks_of_hss = in.begin_blocks_of_hss;
    nr.qnumbers = in.qnumbers;
    int nops = 0;
    for (const auto& i : in.instr) nops = std::max(nops, i.instr[2] + 1);
    nr.nrops = nops;
    nr.hss_of_ops.assign(in.hss_of_new.begin(), in.hss_of_new.begin() + nops);
    nr.beginblock.resize(nops);
    for (int k = 0; k < nops; ++k) nr.beginblock[k] = in.hss_bb[nr.hss_of_ops[k]];
    nr.finalize_offsets();
    std::vector<std::vector<int>> instr_of(nops);
    for (int i = 0; i < ni; ++i) instr_of[in.instr[i].instr[2]].push_back(i);
    for (int k = 0; k < nops; ++k) {
        const int nb = nr.nblocks_op(k);
        for (int b = 0; b < nb; ++b) {
            AppendGroup g;
            g.dst = nr.op_offset[k] + nr.beginblock[k][b];
            g.n = nr.block_size(k, b);
            g.op = k;
            g.block = b;
            g.task_begin = (int)P.stasks.size();
            for (int i : instr_of[k]) {
                const int u = run_of[i];
                P.stasks.push_back(AppendTask{U.op_offset[u] + U.beginblock[u][b], 1.0, in.instr[i].pref});
            }
            g.task_end = (int)P.stasks.size();
            if (g.task_end > g.task_begin && g.n > 0) P.groups.push_back(g);
            else P.stasks.resize(g.task_begin);
        }
    }
    return P;
}

}  // namespace t3
