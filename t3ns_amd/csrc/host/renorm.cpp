// Append plan (src/rOperators_pUpdate.c:393-673, src/rOperators_init.c:185-265).
#include "renorm.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

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
