// Block structures of site tensors and renormalised operators; see tensors.hpp.
#include "tensors.hpp"

#include <algorithm>
#include <cassert>
#include <numeric>
#include <stdexcept>

namespace t3 {

void ROperators::finalize_offsets() {
    op_offset.assign(nrops + 1, 0);
    for (int i = 0; i < nrops; ++i) op_offset[i + 1] = op_offset[i] + beginblock[i].back();
}

namespace {
// Stable permutation sorting qn tuples of width w, last entry most
// significant (src/sort.c:104-133, qsort_r with a stable merge sort).
std::vector<int> sort_qn_tuples(const std::vector<QN>& qn, int w) {
    const int n = w ? (int)(qn.size() / w) : 0;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        for (int k = w - 1; k >= 0; --k) {
            const QN x = qn[(size_t)a * w + k], y = qn[(size_t)b * w + k];
            if (x != y) return x < y;
        }
        return false;
    });
    return idx;
}
}  // namespace

SiteTensor init_1siteTensor(const Context& c, int site) {
    int b[3];
    c.netw.bonds_of_site(site, b);
    const Symsecs* ss[3] = {&c.symsecs(b[0]), &c.symsecs(b[1]), &c.symsecs(b[2])};
    return init_1siteTensor(c, site, ss);
}

SiteTensor init_1siteTensor(const Context& c, int site, const Symsecs* const ss[3]) {
    SiteTensor t;
    t.nrsites = 1;
    t.sites[0] = site;
    GoodSectors gs = c.find_good_sectors(ss, 1);
    std::vector<QN> qn;
    std::vector<int64_t> dims;
    const int n0 = ss[0]->nrSecs, n1 = ss[1]->nrSecs;
    for (int i = 0; i < n0; ++i) {
        if (!gs.rowvalid[i]) continue;
        for (int j = 0; j < n1; ++j)
            for (const GSec& g : gs.sec[i][j]) {
                if (g.d == 0) continue;
                qn.push_back((QN)i + (QN)j * n0 + (QN)g.id3 * n0 * n1);
                dims.push_back(g.d);
            }
    }
    auto idx = sort_qn_tuples(qn, 1);
    t.nrblocks = (int)qn.size();
    t.qnumbers.resize(t.nrblocks);
    t.beginblock.assign(t.nrblocks + 1, 0);
    for (int i = 0; i < t.nrblocks; ++i) {
        t.qnumbers[i] = qn[idx[i]];
        t.beginblock[i + 1] = t.beginblock[i] + dims[idx[i]];
    }
    t.tel.assign(t.size(), 0.0);
    return t;
}

int rOperators_site_to_attach(const Context& c, const ROperators& r) {
    return r.P_operator ? c.netw.bonds[r.bond][!r.is_left] : c.netw.bonds[r.bond][r.is_left];
}

namespace {
GoodSectors good_sectors_mpo(const Context& c, int bond, int is_left) {
    // [MPO, ket or bra, bra or ket] of the bond (src/rOperators_init.c:354-370)
    const Symsecs* ss[3] = {&c.symsecs(-1),
                            &c.symsecs(is_left ? c.get_ketT3NSbond(bond) : c.get_braT3NSbond(bond)),
                            &c.symsecs(!is_left ? c.get_ketT3NSbond(bond) : c.get_braT3NSbond(bond))};
    return c.find_good_sectors(ss, 1);
}
}  // namespace

ROperators init_rOperators_structure(const Context& c, int bond, int is_left, bool P,
                                     std::vector<std::vector<int64_t>>* tmpbb) {
    ROperators r;
    r.bond = bond;
    r.is_left = is_left;
    r.P_operator = P ? 1 : 0;
    r.nrhss = c.ham->nr_hss();
    r.begin_blocks_of_hss.assign(r.nrhss + 1, 0);
    tmpbb->assign(r.nrhss, {});
    GoodSectors gs = good_sectors_mpo(c, bond, is_left);
    const int id0 = 1 + is_left, id1 = 1 + !is_left;
    const int nb0 = gs.ss[id0]->nrSecs, nb1 = gs.ss[id1]->nrSecs;
    if (!P) {
        // src/rOperators_init.c:267-380
        std::vector<std::vector<QN>> qh(r.nrhss);
        for (int h = 0; h < r.nrhss; ++h) {
            auto ents = c.iter_gs(gs, 0, h);
            std::vector<QN> qt;
            std::vector<int64_t> dt;
            for (auto& e : ents) {
                qt.push_back((QN)e.cid[id0] + (QN)e.cid[id1] * nb0 + (QN)h * nb0 * nb1);
                dt.push_back(e.cdim);
            }
            auto idx = sort_qn_tuples(qt, 1);
            auto& bb = (*tmpbb)[h];
            bb.assign(qt.size() + 1, 0);
            qh[h].resize(qt.size());
            for (size_t i = 0; i < qt.size(); ++i) {
                qh[h][i] = qt[idx[i]];
                bb[i + 1] = bb[i] + dt[idx[i]];
            }
            r.begin_blocks_of_hss[h + 1] = r.begin_blocks_of_hss[h] + (int)qt.size();
        }
        for (int h = 0; h < r.nrhss; ++h) r.qnumbers.insert(r.qnumbers.end(), qh[h].begin(), qh[h].end());
        return r;
    }
    // Physical operators (src/rOperators_init.c:383-507).
    if (!c.bond_is_internal_dims1(bond))
        throw std::runtime_error("physical rOperators need an internal bond with unit dims");
    const int site = c.netw.bonds[bond][!is_left];
    int sb[3];
    c.netw.bonds_of_site(site, sb);
    const Symsecs* sss[3] = {&c.symsecs(sb[0]), &c.symsecs(sb[1]), &c.symsecs(sb[2])};
    GoodSectors sitegs = c.find_good_sectors(sss, 1);
    const int ibond = is_left * 2;
    std::vector<std::vector<GsEntry>> qna(sss[ibond]->nrSecs);
    for (int i = 0; i < sss[ibond]->nrSecs; ++i) qna[i] = c.iter_gs(sitegs, ibond, i);
    std::vector<std::vector<QN>> qh(r.nrhss);
    for (int h = 0; h < r.nrhss; ++h) {
        std::vector<QN> qt;
        std::vector<int64_t> dt;
        for (auto& e : c.iter_gs(gs, 0, h)) {
            const int bra = e.cid[id0], ket = e.cid[id1];
            const QN mqn = (QN)bra + (QN)ket * nb0 + (QN)h * nb0 * nb1;
            for (auto& qb : qna[bra])
                for (auto& qk : qna[ket]) {
                    qt.push_back(qb.cqn);
                    qt.push_back(qk.cqn);
                    qt.push_back(mqn);
                    dt.push_back((int64_t)qb.cdim * qk.cdim);
                }
        }
        auto idx = sort_qn_tuples(qt, 3);
        auto& bb = (*tmpbb)[h];
        bb.assign(dt.size() + 1, 0);
        qh[h].resize(qt.size());
        for (size_t i = 0; i < dt.size(); ++i) {
            for (int k = 0; k < 3; ++k) qh[h][i * 3 + k] = qt[(size_t)idx[i] * 3 + k];
            bb[i + 1] = bb[i] + dt[idx[i]];
        }
        r.begin_blocks_of_hss[h + 1] = r.begin_blocks_of_hss[h] + (int)dt.size();
    }
    for (int h = 0; h < r.nrhss; ++h) r.qnumbers.insert(r.qnumbers.end(), qh[h].begin(), qh[h].end());
    return r;
}

ROperators vacuum_rOperators(const Context& c, int bond, int is_left) {
    if (c.netw.bonds[bond][!is_left] != -1) throw std::runtime_error("not a border bond");
    ROperators r;
    r.bond = bond;
    r.is_left = is_left;
    r.P_operator = 0;
    r.nrhss = c.ham->nr_hss();
    const int triv = c.ham->trivial_hss();
    const Symsecs& ss = c.symsecs(bond);
    const int n = ss.nrSecs;
    const int nb = is_left ? n : n * n;
    r.begin_blocks_of_hss.assign(r.nrhss + 1, 0);
    for (int i = triv + 1; i <= r.nrhss; ++i) r.begin_blocks_of_hss[i] = nb;
    for (int i = 0; i < nb; ++i)
        r.qnumbers.push_back(is_left ? (QN)i * (n + 1) + (QN)n * n * triv : (QN)i + (QN)n * n * triv);
    r.nrops = 1;
    r.hss_of_ops = {triv};
    // Unit operator: diagonal, mirrored coupling prefactor for right
    // operators (src/rOperators_init.c:59-110).
    const Symsecs& mpo = c.symsecs(-1);
    std::vector<int64_t> bb(nb + 1, 0);
    std::vector<int> D(nb);
    for (int b = 0; b < nb; ++b) {
        QN q = r.qnumbers[b];
        const int i0 = (int)(q % n);
        q /= n;
        const int i1 = (int)(q % n);
        const int i2 = (int)(q / n);
        D[b] = ss.dims[i0] * ss.dims[i1] * mpo.dims[i2];
        bb[b + 1] = bb[b] + (int64_t)D[b] * D[b];
    }
    r.beginblock = {bb};
    r.finalize_offsets();
    r.tel.assign(r.total_size(), 0.0);
    for (int b = 0; b < nb; ++b) {
        QN q = r.qnumbers[b];
        const int i0 = (int)(q % n);
        q /= n;
        const int i1 = (int)(q % n);
        const int i2 = (int)(q / n);
        const int* irr[3] = {ss.irreps[i0].data(), ss.irreps[i1].data(), mpo.irreps[i2].data()};
        const double pref = is_left ? 1.0 : prefactor_mirror_coupling(irr, c.sym);
        for (int j = 0; j < D[b]; ++j) r.tel[bb[b] + (int64_t)j * D[b] + j] = pref;
    }
    return r;
}

std::vector<int> internal_bonds(const Context& c, const SiteTensor& t) {
    std::vector<int> bonds(3 * t.nrsites), out;
    for (int i = 0; i < t.nrsites; ++i) c.netw.bonds_of_site(t.sites[i], &bonds[3 * i]);
    for (int i = 0; i < 3 * t.nrsites; ++i)
        for (int j = i + 1; j < 3 * t.nrsites; ++j)
            if (bonds[i] == bonds[j]) {
                out.push_back(c.get_ketT3NSbond(bonds[i]));
                break;
            }
    return out;
}

SiteTensor make_multisite_structure(Context& c, const int* sitelist, int nr, MultisiteInfo* md) {
    // init_md (src/siteTensor_init.c:180-260)
    SiteTensor T;
    T.nrsites = 0;
    md->nr_internal = 0;
    for (auto& cc : md->connected) cc[0] = cc[1] = -1;
    int branch = -1;
    for (int i = 0; i < nr; ++i) {
        if (c.netw.is_psite(sitelist[i])) continue;
        if (branch != -1) throw std::runtime_error("multisite tensor with two branching sites");
        branch = sitelist[i];
    }
    auto add_psite = [&](int bond, int bid) {
        const int site = c.netw.bonds[bond][bid == 2];
        if (std::find(sitelist, sitelist + nr, site) == sitelist + nr) {
            md->connected[bid][0] = -1;
            md->connected[bid][1] = -1;
        } else {
            md->connected[bid][0] = T.nrsites;
            md->connected[bid][1] = 2 * (bid != 2);
            md->outersites[md->nr_internal] = T.nrsites;
            md->internals[md->nr_internal] = bond;
            T.sites[T.nrsites++] = site;
            ++md->nr_internal;
        }
    };
    if (branch != -1) {
        int b[3];
        c.netw.bonds_of_site(branch, b);
        add_psite(b[0], 0);
        add_psite(b[1], 1);
        md->innersite = T.nrsites;
        T.sites[T.nrsites++] = branch;
        add_psite(b[2], 2);
    } else {
        if (nr != 2) throw std::runtime_error("multisite tensor without branch needs 2 sites");
        const int bond = c.netw.common_bond(sitelist[0], sitelist[1]);
        T.nrsites = 2;
        md->nr_internal = 1;
        md->internals[0] = bond;
        T.sites[0] = c.netw.bonds[bond][0];
        T.sites[1] = c.netw.bonds[bond][1];
        md->outersites[0] = 0;
        md->outersites[1] = 1;
        md->outersites[2] = 2;
        md->innersite = nr - 1;
        md->connected[0][0] = 0;
        md->connected[0][1] = 2;
    }
    if (md->nr_internal != nr - 1) throw std::runtime_error("multisite: bad site list");
    for (int i = 0; i < T.nrsites; ++i) {
        c.netw.bonds_of_site(T.sites[i], md->bonds[i]);
        for (int j = 0; j < 3; ++j) md->ssarr_old[i][j] = c.symsecs(md->bonds[i][j]);
    }

    // make_internalss_and_tensor (src/siteTensor_init.c:609-646)
    auto ib = internal_bonds(c, T);
    md->nr_internal = (int)ib.size();
    for (int i = 0; i < md->nr_internal; ++i) md->internals[i] = ib[i];
    std::vector<Symsecs> intss(md->nr_internal);
    for (int i = 0; i < md->nr_internal; ++i) {
        // initialize_inner_symsecs (src/siteTensor_init.c:270-297)
        const int site = md->outersites[i];
        const int* bonds = md->bonds[site];
        const int sign = md->internals[i] == bonds[2] ? 1 : -1;
        const Symsecs* s[2];
        int cnt = 0;
        for (int j = 2; j >= 0; --j)
            if (bonds[j] != md->internals[i]) s[cnt++] = &md->ssarr_old[site][j];
        intss[i] = c.tensprod_symsecs(*s[0], *s[1], sign, 'n');
        intss[i].bond = md->internals[i];
    }
    auto set_ssarr = [&]() {
        for (int i = 0; i < T.nrsites; ++i)
            for (int j = 0; j < 3; ++j) {
                int k;
                for (k = 0; k < md->nr_internal; ++k)
                    if (md->bonds[i][j] == md->internals[k]) break;
                md->ssarr[i][j] = k == md->nr_internal ? md->ssarr_old[i][j] : intss[k];
            }
    };
    set_ssarr();
    {
        // correct_inner_symsecs (src/siteTensor_init.c:299-335)
        const int* bonds = md->bonds[md->innersite];
        const Symsecs* ss[3] = {&md->ssarr[md->innersite][0], &md->ssarr[md->innersite][1],
                                &md->ssarr[md->innersite][2]};
        GoodSectors gs = c.find_good_sectors(ss, 1);
        for (int i = 0; i < 3; ++i) {
            int j;
            for (j = 0; j < md->nr_internal; ++j)
                if (bonds[i] == md->internals[j]) break;
            if (j == md->nr_internal) continue;
            for (int s = 0; s < intss[j].nrSecs; ++s)
                if (c.iter_gs_length(gs, i, s) < 1) {
                    intss[j].fcidims[s] = 0;
                    intss[j].dims[s] = 0;
                }
        }
        for (auto& s : intss) c.kick_empty_symsecs(s, 'n');
    }
    set_ssarr();

    // init_multisitetensor (src/siteTensor_init.c:337-560)
    std::vector<GoodSectors> gs(T.nrsites);
    for (int i = 0; i < T.nrsites; ++i) {
        const Symsecs* ss[3] = {&md->ssarr[i][0], &md->ssarr[i][1], &md->ssarr[i][2]};
        gs[i] = c.find_good_sectors(ss, 1);
    }
    std::vector<QN> qn;
    std::vector<int64_t> dims;
    const int isite = md->innersite;
    const Symsecs* iss = md->ssarr[isite];
    for (int i = 0; i < iss[0].nrSecs; ++i) {
        if (!gs[isite].rowvalid[i]) continue;
        for (int j = 0; j < iss[1].nrSecs; ++j) {
            const auto& gsa = gs[isite].sec[i][j];
            for (size_t k = 0; k < gsa.size(); ++k) {
                const int ids[3] = {i, j, gsa[k].id3};
                std::vector<std::vector<GsEntry>> part(T.nrsites);
                bool empty = false;
                for (int leg = 0; leg < 3; ++leg) {
                    if (md->connected[leg][0] == -1) continue;
                    part[md->connected[leg][0]] =
                        c.iter_gs(gs[md->connected[leg][0]], md->connected[leg][1], ids[leg]);
                    if (part[md->connected[leg][0]].empty()) empty = true;
                }
                if (empty) continue;
                GsEntry inner;
                inner.cqn = (QN)ids[0] + (QN)ids[1] * iss[0].nrSecs +
                            (QN)ids[2] * iss[0].nrSecs * iss[1].nrSecs;
                inner.cdim = gsa[k].d;
                part[isite] = {inner};
                std::vector<int> ix(T.nrsites, 0);
                while (true) {
                    int64_t d = 1;
                    for (int s = 0; s < T.nrsites; ++s) {
                        qn.push_back(part[s][ix[s]].cqn);
                        d *= part[s][ix[s]].cdim;
                    }
                    dims.push_back(d);
                    int s;
                    for (s = 0; s < T.nrsites; ++s) {
                        if (++ix[s] < (int)part[s].size()) break;
                        ix[s] = 0;
                    }
                    if (s == T.nrsites) break;
                }
            }
        }
    }
    // sort_and_make (src/siteTensor_init.c:337-360)
    const int ns = T.nrsites;
    auto idx = sort_qn_tuples(qn, ns);
    T.nrblocks = (int)dims.size();
    T.qnumbers.resize(qn.size());
    T.beginblock.assign(T.nrblocks + 1, 0);
    for (int i = 0; i < T.nrblocks; ++i) {
        for (int j = 0; j < ns; ++j) T.qnumbers[(size_t)i * ns + j] = qn[(size_t)idx[i] * ns + j];
        T.beginblock[i + 1] = T.beginblock[i] + dims[idx[i]];
    }
    // Install the internal sectors (src/siteTensor_init.c:262-268).
    for (int i = 0; i < md->nr_internal; ++i) c.v_symsecs[md->internals[i]] = intss[i];
    return T;
}

}  // namespace t3
