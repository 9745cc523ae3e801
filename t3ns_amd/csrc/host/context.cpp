// Symmetry-sector bookkeeping; see context.hpp.
// Follows src/bookkeeper.c:32-330, src/tensorproducts.c, src/symsecs.c and
// src/instructions.c:51-251.
#include "context.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace t3 {

// ------------------------------------------------------------------ lookup
const Symsecs& Context::symsecs(int bond) const {
    if (bond >= 2 * netw.nr_bonds) return p_symsecs[(bond - 2 * netw.nr_bonds) % netw.psites];
    if (bond >= 0) return v_symsecs[bond % netw.nr_bonds];
    if (bond == -1) return ham->mpo_symsecs();
    throw std::runtime_error("symsecs: bad bond");
}
Symsecs& Context::symsecs_mut(int bond) {
    if (bond >= 2 * netw.nr_bonds) return p_symsecs[(bond - 2 * netw.nr_bonds) % netw.psites];
    if (bond >= 0) return v_symsecs[bond % netw.nr_bonds];
    throw std::runtime_error("symsecs_mut: bad bond");
}
int Context::get_braT3NSbond(int bond) const {
    if (bond < netw.nr_bonds) return bond + netw.nr_bonds;
    return bond + netw.psites;
}
int Context::get_ketT3NSbond(int bond) const { return bond; }

bool Context::irreps_equal(const int* a, const int* b) const {
    for (int i = 0; i < sym.nrSyms; ++i)
        if (a[i] != b[i]) return false;
    return true;
}

namespace {
// Order of sector labels: last symmetry most significant (src/sort.c:36-75).
int cmp_irrep(const int* a, const int* b, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
}  // namespace

int Context::search_symsec(const int* irrep, const Symsecs& ss) const {
    // Virtual bonds are sorted: bisection; physical and MPO bonds: linear
    // scan (src/symsecs.c:164-178).
    const bool sorted = ss.bond >= 0 && !netw.is_pbond(ss.bond);
    if (sorted) {
        int lo = 0, hi = ss.nrSecs - 1;
        while (lo <= hi) {
            const int mid = (lo + hi) / 2;
            const int c = cmp_irrep(irrep, ss.irreps[mid].data(), sym.nrSyms);
            if (c == 0) return mid;
            if (c < 0) hi = mid - 1;
            else lo = mid + 1;
        }
        return -1;
    }
    for (int i = 0; i < ss.nrSecs; ++i)
        if (irreps_equal(irrep, ss.irreps[i].data())) return i;
    return -1;
}

// --------------------------------------------------------- sector algebra
void Context::build_all_sectors(Symsecs& res, const Symsecs& s1, const Symsecs& s2) const {
    // Direct product of the label ranges, filtered by the model's
    // consistency rule (src/tensorproducts.c:251-302).
    int maxirr[MAX_SYMMETRIES], idx[MAX_SYMMETRIES] = {0};
    long total = 1;
    for (int i = 0; i < sym.nrSyms; ++i) {
        int m1 = 0, m2 = 0;
        for (int k = 0; k < s1.nrSecs; ++k) m1 = std::max(m1, s1.irreps[k][i]);
        for (int k = 0; k < s2.nrSecs; ++k) m2 = std::max(m2, s2.irreps[k][i]);
        maxirr[i] = max_irrep_bound(m1, m2, sym.sgs[i]);
        total *= maxirr[i];
    }
    res.irreps.clear();
    for (long c = 0; c < total; ++c) {
        if (ham->consistent_state(idx, sym)) {
            Irrep ir{};
            for (int i = 0; i < sym.nrSyms; ++i) ir[i] = idx[i];
            res.irreps.push_back(ir);
        }
        for (int i = 0; i < sym.nrSyms; ++i) {
            if (++idx[i] == maxirr[i]) idx[i] = 0;
            else break;
        }
    }
    res.nrSecs = (int)res.irreps.size();
}

void Context::kick_empty_symsecs(Symsecs& s, char o) const {
    // src/symsecs.c:96-126
    int cnt = 0;
    if (o == 'n') s.totaldims = 0;
    for (int i = 0; i < s.nrSecs; ++i) {
        if (s.fcidims[i] < 0.5 || (o == 'n' && s.dims[i] == 0)) continue;
        s.irreps[cnt] = s.irreps[i];
        s.fcidims[cnt] = s.fcidims[i];
        if (o == 'n') {
            s.dims[cnt] = s.dims[i];
            s.totaldims += s.dims[cnt];
        }
        ++cnt;
    }
    s.nrSecs = cnt;
    s.irreps.resize(cnt);
    s.fcidims.resize(cnt);
    if (o == 'n') s.dims.resize(cnt);
}

Symsecs Context::tensprod_symsecs(const Symsecs& s1, const Symsecs& s2, int sign, char o) const {
    // src/tensorproducts.c:337-414
    Symsecs res;
    build_all_sectors(res, s1, s2);
    res.fcidims.assign(res.nrSecs, 0.0);
    if (o != 'f') res.dims.assign(res.nrSecs, 0);
    auto zero = [o](const Symsecs& s, int i) {
        return o == 'f' ? s.fcidims[i] < 0.5 : s.dims[i] == 0;
    };
    for (int i = 0; i < s1.nrSecs; ++i) {
        if (zero(s1, i)) continue;
        for (int j = 0; j < s2.nrSecs; ++j) {
            if (zero(s2, j)) continue;
            int mn[MAX_SYMMETRIES], nr[MAX_SYMMETRIES], st[MAX_SYMMETRIES], cur[MAX_SYMMETRIES];
            int tot = 1;
            for (int y = 0; y < sym.nrSyms; ++y) {
                tensprod_irrep(&mn[y], &nr[y], &st[y], s1.irreps[i][y], s2.irreps[j][y], sign,
                               sym.sgs[y]);
                cur[y] = mn[y];
                tot *= nr[y];
            }
            for (int c = 0; c < tot; ++c) {
                const int ps = search_symsec(cur, res);
                if (ps >= 0) {
                    switch (o) {
                        case 'd':
                            res.dims[ps] += s1.dims[i] * s2.dims[j];
                            res.fcidims[ps] += s1.fcidims[i] * s2.fcidims[j];
                            break;
                        case 'f': res.fcidims[ps] += s1.fcidims[i] * s2.fcidims[j]; break;
                        case 'n':
                            res.fcidims[ps] = 1;
                            res.dims[ps] = 1;
                            break;
                    }
                }
                for (int y = 0; y < sym.nrSyms; ++y) {
                    cur[y] += st[y];
                    if (cur[y] <= mn[y] + (nr[y] - 1) * st[y]) break;
                    cur[y] = mn[y];
                }
            }
        }
    }
    kick_empty_symsecs(res, o == 'd' ? 'n' : o);
    return res;
}

void Context::kick_impossibles(Symsecs& s) const {
    // src/bookkeeper.c:32-69
    int cnt = 0;
    for (int i = 0; i < s.nrSecs; ++i) {
        bool bad = false;
        for (int j = 0; j < sym.nrSyms; ++j)
            if (sym.sgs[j] == U1 && s.irreps[i][j] > sym.target_state[j]) bad = true;
        if (bad) continue;
        s.irreps[cnt] = s.irreps[i];
        s.fcidims[cnt] = s.fcidims[i];
        if (!s.dims.empty()) s.dims[cnt] = s.dims[i];
        ++cnt;
    }
    s.nrSecs = cnt;
    s.irreps.resize(cnt);
    s.fcidims.resize(cnt);
    if (!s.dims.empty()) s.dims.resize(cnt);
}

Symsecs Context::init_vacuum(char o) const {
    Symsecs s;
    s.nrSecs = 1;
    s.irreps.assign(1, Irrep{});
    s.fcidims.assign(1, 1.0);
    if (o == 'd') s.dims.assign(1, 1);
    return s;
}

Symsecs Context::init_target(char o) const {
    Symsecs s;
    s.nrSecs = 1;
    Irrep ir{};
    for (int i = 0; i < sym.nrSyms; ++i) ir[i] = sym.target_state[i];
    s.irreps.assign(1, ir);
    s.fcidims.assign(1, 1.0);
    if (o == 'd') s.dims.assign(1, 1);
    return s;
}

bool Context::select_lowest(Symsecs& s1, const Symsecs& s2) const {
    // src/bookkeeper.c:154-183.  Returns true when nothing had to change.
    bool ok = true;
    for (int i = 0; i < s1.nrSecs; ++i) {
        int j;
        for (j = 0; j < s2.nrSecs; ++j)
            if (irreps_equal(s1.irreps[i].data(), s2.irreps[j].data())) break;
        if (j == s2.nrSecs) {
            s1.fcidims[i] = 0;
            ok = false;
        } else {
            const bool unchanged = s1.fcidims[i] <= s2.fcidims[j];
            ok = ok && unchanged;
            if (!unchanged) s1.fcidims[i] = s2.fcidims[j];
        }
    }
    if (!ok) kick_empty_symsecs(s1, 'f');
    return ok;
}

void Context::make_symsec(Symsecs* symsec, int bond, int is_left, char o) {
    // src/bookkeeper.c:185-253
    if (netw.bonds[bond][!is_left] == -1) {
        if (is_left) {
            *symsec = init_vacuum(o);
        } else {
            Symsecs tmp = init_target(o);
            select_lowest(*symsec, tmp);
        }
        return;
    }
    const int site = netw.bonds[bond][!is_left];
    int i;
    for (i = 0; i < netw.nr_bonds; ++i)
        if (netw.bonds[i][is_left] == site) break;
    const Symsecs s1 = v_symsecs[i];  // value copy (sectors of the neighbour bond)
    Symsecs* s2p;
    const bool psite = netw.is_psite(site);
    if (psite) {
        s2p = &p_symsecs[netw.sitetoorb[site]];
    } else {
        for (i = is_left ? i + 1 : 0; i < netw.nr_bonds; ++i)
            if (netw.bonds[i][1] == site && i != bond) break;
        s2p = &v_symsecs[i];
    }
    if (is_left) {
        *symsec = tensprod_symsecs(s1, *s2p, +1, o);
        kick_impossibles(*symsec);
        symsec->bond = bond;
        return;
    }
    if (psite) {
        Symsecs tmp = tensprod_symsecs(s1, *s2p, -1, o);
        select_lowest(*symsec, tmp);
        symsec->bond = bond;
        return;
    }
    bool flag = false;
    while (!flag) {
        Symsecs tmp = tensprod_symsecs(s1, *s2p, -1, o);
        Symsecs tmp2 = tensprod_symsecs(s1, *symsec, -1, o);
        const bool f1 = select_lowest(*s2p, tmp2);
        const bool f2 = select_lowest(*symsec, tmp);
        flag = f1 && f2;
    }
}

void Context::prepare_bookkeeper(int max_dim, bool interm_scale, int minocc) {
    // Physical sectors (src/bookkeeper.c:322-330).
    p_symsecs.clear();
    for (int i = 0; i < netw.psites; ++i) {
        p_symsecs.push_back(ham->phys_symsecs(i));
        p_symsecs.back().bond = netw.nr_bonds * 2 + i;
    }
    // FCI dimensions, left then right pass (src/bookkeeper.c:255-265).
    v_symsecs.assign(netw.nr_bonds, Symsecs());
    for (int b = 0; b < netw.nr_bonds; ++b) make_symsec(&v_symsecs[b], b, 1, 'f');
    for (int b = netw.nr_bonds - 1; b >= 0; --b) make_symsec(&v_symsecs[b], b, 0, 'f');
    if (!interm_scale) {
        // src/bookkeeper.c:267-289
        for (auto& s : v_symsecs) {
            double tot = 0;
            for (int i = 0; i < s.nrSecs; ++i) tot += s.fcidims[i];
            const double ratio = max_dim < tot ? max_dim * 1. / tot : 1;
            s.dims.assign(s.nrSecs, 0);
            s.totaldims = 0;
            for (int i = 0; i < s.nrSecs; ++i) {
                const int old = s.dims[i];
                s.dims[i] = (int)std::ceil(ratio * s.fcidims[i]);
                if (s.dims[i] < minocc) s.dims[i] = old < minocc ? old : minocc;
                s.totaldims += s.dims[i];
            }
        }
        return;
    }
    // Intermediate scaling (src/bookkeeper.c:291-320).
    for (int b = 0; b < netw.nr_bonds; ++b) {
        Symsecs ns;
        make_symsec(&ns, b, 1, 'd');
        Symsecs& bs = v_symsecs[b];
        bs.dims.assign(bs.nrSecs, 0);
        bs.totaldims = 0;
        for (int i = 0; i < bs.nrSecs; ++i) {
            const int pos = search_symsec(bs.irreps[i].data(), ns);
            if (pos < 0) throw std::runtime_error("prepare_bookkeeper: irrep not found");
            bs.dims[i] = ns.dims[pos] > bs.fcidims[i] ? (int)bs.fcidims[i] : ns.dims[pos];
            bs.totaldims += bs.dims[i];
        }
        if (bs.totaldims <= max_dim) continue;
        const double ratio = max_dim * 1. / bs.totaldims;
        bs.totaldims = 0;
        for (int i = 0; i < bs.nrSecs; ++i) {
            const int old = bs.dims[i];
            bs.dims[i] = (int)std::ceil(ratio * bs.dims[i]);
            if (bs.dims[i] < minocc) bs.dims[i] = old < minocc ? old : minocc;
            bs.totaldims += bs.dims[i];
        }
    }
    for (int b = 0; b < netw.nr_bonds; ++b) v_symsecs[b].bond = b;
}

GoodSectors Context::find_good_sectors(const Symsecs* const ss[3], int sign) const {
    // src/tensorproducts.c:66-142
    GoodSectors gs;
    for (int k = 0; k < 3; ++k) gs.ss[k] = ss[k];
    const int n0 = ss[0]->nrSecs, n1 = ss[1]->nrSecs;
    gs.rowvalid.assign(n0, 0);
    gs.sec.assign(n0, {});
    for (int i = 0; i < n0; ++i) {
        if (ss[0]->dims[i] == 0) continue;
        gs.rowvalid[i] = 1;
        gs.sec[i].assign(n1, {});
        for (int j = 0; j < n1; ++j) {
            if (ss[1]->dims[j] == 0) continue;
            const int dim = ss[0]->dims[i] * ss[1]->dims[j];
            int mn[MAX_SYMMETRIES], nr[MAX_SYMMETRIES], st[MAX_SYMMETRIES], cur[MAX_SYMMETRIES];
            int tot = 1;
            for (int y = 0; y < sym.nrSyms; ++y) {
                tensprod_irrep(&mn[y], &nr[y], &st[y], ss[0]->irreps[i][y], ss[1]->irreps[j][y],
                               sign, sym.sgs[y]);
                cur[y] = mn[y];
                tot *= nr[y];
            }
            for (int c = 0; c < tot; ++c) {
                const int ind = search_symsec(cur, *ss[2]);
                if (ind != -1 && ss[2]->dims[ind] != 0) {
                    gs.sec[i][j].push_back(GSec{dim * ss[2]->dims[ind], ind});
                    ++gs.total;
                }
                for (int y = 0; y < sym.nrSyms; ++y) {
                    cur[y] += st[y];
                    if (cur[y] <= mn[y] + (nr[y] - 1) * st[y]) break;
                    cur[y] = mn[y];
                }
            }
        }
    }
    return gs;
}

std::vector<GsEntry> Context::iter_gs(const GoodSectors& gs, int idnr, int tid) const {
    std::vector<GsEntry> out;
    const int n0 = gs.ss[0]->nrSecs, n1 = gs.ss[1]->nrSecs;
    auto push = [&](int i, int j, int k3, int d) {
        GsEntry e;
        e.cid[0] = i;
        e.cid[1] = j;
        e.cid[2] = k3;
        e.cqn = (QN)i + (QN)j * n0 + (QN)k3 * n0 * n1;
        e.cdim = d;
        out.push_back(e);
    };
    if (idnr == 2) {
        for (int i = 0; i < n0; ++i) {
            if (!gs.rowvalid[i]) continue;
            for (int j = 0; j < n1; ++j) {
                for (const GSec& g : gs.sec[i][j])
                    if (g.id3 == tid) {
                        if (g.d != 0) push(i, j, tid, g.d);
                        break;
                    }
            }
        }
    } else if (idnr == 0) {
        if (gs.rowvalid[tid])
            for (int j = 0; j < n1; ++j)
                for (const GSec& g : gs.sec[tid][j])
                    if (g.d != 0) push(tid, j, g.id3, g.d);
    } else {
        for (int i = 0; i < n0; ++i) {
            if (!gs.rowvalid[i]) continue;
            for (const GSec& g : gs.sec[i][tid])
                if (g.d != 0) push(i, tid, g.id3, g.d);
        }
    }
    return out;
}

int Context::iter_gs_length(const GoodSectors& gs, int idnr, int tid) const {
    return (int)iter_gs(gs, idnr, tid).size();
}

bool Context::bond_is_internal_dims1(int bond) const {
    const Symsecs& s = symsecs(get_ketT3NSbond(bond));
    for (int i = 0; i < s.nrSecs; ++i)
        if (s.dims[i] != 1) return false;
    return true;
}

int Context::full_dimension(const Symsecs& s) const {
    if (!need_multiplicity(sym)) return s.totaldims;
    int r = 0;
    for (int i = 0; i < s.nrSecs; ++i) r += s.dims[i] * multiplicity(sym, s.irreps[i].data());
    return r;
}

// ------------------------------------------------------------ instructions
void sort_instructions(InstructionSet& s) {
    std::stable_sort(s.instr.begin(), s.instr.end(), [](const Instruction& a, const Instruction& b) {
        if (a.instr[0] != b.instr[0]) return a.instr[0] < b.instr[0];
        return a.instr[1] < b.instr[1];
    });
}

void sort_instructions_merge(InstructionSet& s, const int* const* hss_ops, int nrhss) {
    const int n = s.nr_instr();
    std::vector<int> key(n), idx(n);
    for (int i = 0; i < n; ++i) {
        int t = 0;
        for (int j = s.step - 1; j >= 0; --j) t = t * nrhss + hss_ops[j][s.instr[i].instr[j]];
        key[i] = t;
    }
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<Instruction> ni(n);
    s.MPOc.clear();
    s.MPOc_beg.clear();
    for (int i = 0; i < n; ++i) {
        ni[i] = s.instr[idx[i]];
        if (i == 0 || s.MPOc.back() != key[idx[i]]) {
            s.MPOc_beg.push_back(i);
            s.MPOc.push_back(key[idx[i]]);
        }
    }
    s.MPOc_beg.push_back(n);
    s.instr = std::move(ni);
}

bool next_unique_instr(int* curr, const InstructionSet& s) {
    if (*curr == -1) {
        ++*curr;
        return s.nr_instr() > 0;
    }
    const int o1 = s.instr[*curr].instr[0], o2 = s.instr[*curr].instr[1];
    const int ho = s.hss_of_new[s.instr[*curr].instr[2]];
    for (++*curr; *curr < s.nr_instr(); ++*curr) {
        const Instruction& c = s.instr[*curr];
        if (c.instr[0] != o1 || c.instr[1] != o2 || s.hss_of_new[c.instr[2]] != ho) return true;
    }
    return false;
}

const InstructionSet& Context::fetch_pUpdate(int bond, int is_left) {
    auto key = std::make_pair(bond, is_left);
    auto it = pupd_.find(key);
    if (it != pupd_.end()) return it->second;
    InstructionSet s = ham->raw_pUpdate(netw, bond, is_left);
    sort_instructions(s);
    return pupd_[key] = std::move(s);
}

const InstructionSet& Context::fetch_bUpdate(int bond, int is_left) {
    auto key = std::make_pair(bond, is_left);
    auto it = bupd_.find(key);
    if (it != bupd_.end()) return it->second;
    InstructionSet s = ham->raw_bUpdate(netw, bond, is_left);
    sort_instructions(s);
    return bupd_[key] = std::move(s);
}

const InstructionSet& Context::fetch_merge(int bond, int isdmrg, const int* const* hss_ops) {
    auto key = std::make_pair(bond, isdmrg);
    auto it = merge_.find(key);
    if (it != merge_.end()) return it->second;
    InstructionSet s = ham->raw_merge(netw, bond, isdmrg);
    sort_instructions_merge(s, hss_ops, ham->nr_hss());
    return merge_[key] = std::move(s);
}

const Context::RawMerge& Context::fetch_raw_merge(int bond, int isdmrg) {
    auto key = std::make_pair(bond, isdmrg);
    auto it = rawmerge_.find(key);
    if (it != rawmerge_.end()) return it->second;
    RawMerge m;
    m.set = ham->raw_merge(netw, bond, isdmrg);
    m.instr.reserve(3 * m.set.instr.size());
    m.pref.reserve(m.set.instr.size());
    for (const auto& ins : m.set.instr) {
        for (int k = 0; k < 3; ++k) m.instr.push_back(ins.instr[k]);
        m.pref.push_back(ins.pref);
    }
    return rawmerge_[key] = std::move(m);
}

std::unique_ptr<Context> make_context(Network netw, std::unique_ptr<Model> ham,
                                      const std::vector<int>& sgs_in,
                                      const std::vector<int>& ts_in) {
    if (sgs_in.size() != ts_in.size()) throw std::runtime_error("symmetries/target size mismatch");
    // Sort symmetries by group and permute the target state accordingly
    // (src/io.c:246-323), then prepend Z2 if absent (src/bookkeeper.c:495-513).
    std::vector<int> order(sgs_in.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sgs_in[a] < sgs_in[b]; });
    auto ctx = std::make_unique<Context>();
    SymInfo& s = ctx->sym;
    s.nrSyms = 0;
    if (sgs_in.empty() || sgs_in[order[0]] != Z2) {
        s.sgs[s.nrSyms] = Z2;
        s.target_state[s.nrSyms] = 0;
        ++s.nrSyms;
    }
    for (int o : order) {
        if (s.nrSyms >= MAX_SYMMETRIES) throw std::runtime_error("too many symmetries");
        s.sgs[s.nrSyms] = sgs_in[o];
        s.target_state[s.nrSyms] = ts_in[o];
        ++s.nrSyms;
    }
    if (sgs_in.empty() || sgs_in[order[0]] != Z2) {
        // Parity from the U1 labels, else from SU2 (src/bookkeeper.c:462-493).
        int z = 0;
        bool found = false;
        for (int i = 1; i < s.nrSyms; ++i)
            if (s.sgs[i] == U1) {
                z += s.target_state[i];
                found = true;
            }
        if (!found)
            for (int i = 1; i < s.nrSyms; ++i)
                if (s.sgs[i] == SU2) {
                    z += s.target_state[i];
                    found = true;
                }
        if (!found) throw std::runtime_error("symmetries do not fix the Z2 parity");
        s.target_state[0] = z % 2;
    }
    ctx->netw = std::move(netw);
    ctx->ham = std::move(ham);
    ctx->ham->attach(ctx->netw, ctx->sym);
    return ctx;
}

}  // namespace t3
