// C ABI of the host layer (include/t3ns_host.h): contexts and steps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/t3ns_host.h"
#include "host/checkpoint.hpp"
#include "host/context.hpp"
#include "host/tensors.hpp"

using namespace t3;

namespace t3 {
void set_last_error(const std::string& msg);
}

namespace {
template <typename F>
int guarded_host(F&& f) {
    try {
        f();
        return T3NS_OK;
    } catch (const std::bad_alloc& e) {
        t3::set_last_error(e.what());
        return T3NS_ERR_NOMEM;
    } catch (const std::exception& e) {
        t3::set_last_error(e.what());
        return T3NS_ERR_INVALID;
    }
}

// SplitMix64 + Box-Muller: a small deterministic normal generator.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

// Flattened copies of symsecs for descriptors.
struct SymsecsFlat {
    std::vector<int> irreps, dims;
    t3ns_symsecs_desc desc{};
    void set(const Symsecs& s) {
        irreps.assign((size_t)s.nrSecs * T3NS_MAX_SYMMETRIES, 0);
        for (int i = 0; i < s.nrSecs; ++i)
            for (int k = 0; k < T3NS_MAX_SYMMETRIES; ++k) irreps[i * T3NS_MAX_SYMMETRIES + k] = s.irreps[i][k];
        dims = s.dims;
        desc.bond = s.bond;
        desc.nrSecs = s.nrSecs;
        desc.irreps = irreps.data();
        desc.dims = dims.data();
    }
};

// nrops and MPO sector of every operator at `bond` as produced by the
// renormalisation that creates it (vacuum, physical or branching update).
void operator_types(Context& c, int bond, int is_left, int* nrops, std::vector<int>* hss) {
    const int far = c.netw.bonds[bond][!is_left];
    if (far == -1) {
        *nrops = 1;
        *hss = {c.ham->trivial_hss()};
        return;
    }
    const InstructionSet* s;
    if (c.netw.is_psite(far)) {
        int b[3];
        c.netw.bonds_of_site(far, b);
        s = &c.fetch_pUpdate(b[is_left ? 0 : 2], is_left);
    } else {
        s = &c.fetch_bUpdate(bond, is_left);
    }
    int n = 0;
    for (const auto& i : s->instr) n = std::max(n, i.instr[2] + 1);
    *nrops = n;
    hss->assign(s->hss_of_new.begin(), s->hss_of_new.begin() + n);
}
}  // namespace

// An update_rOperators_branching problem: seeded old operators on two bonds
// of a branching site, a seeded branching tensor and the flattened inputs.
struct BUpdateStore {
    ROperators ops[2];
    std::vector<const int64_t*> bb_ptr[2];
    std::vector<const double*> tel_ptr[2];
    SymsecsFlat site_ss[3], mpo;
    SiteTensor T;
    std::vector<int> instr, hss_of_new, tprod_begin, tprods, bbh;
    std::vector<double> pref;
    std::vector<QN> qnumbers;
    std::vector<std::vector<int64_t>> hss_bb;
    std::vector<const int64_t*> hss_bb_ptr;
    int update_case = -1, nrhss = 0;
};

struct QRStore {
    SymsecsFlat ssA[3], ssC[3];
    SiteTensor A, C;
};


// An append problem of one physical leg of a step: the operators of the
// far bond (old) and the flattened rOperators_append_phys inputs.
struct AppendStore {
    ROperators old;
    std::vector<std::vector<int64_t>> old_bb;
    std::vector<const int64_t*> old_bb_ptr;
    std::vector<const double*> old_tel_ptr;
    SymsecsFlat site_ss[3], new_ss[3], old_ss[3], mpo;
    std::vector<int> instr, hss_of_new, siteop_hss, tprod_begin, tprods, bbh;
    std::vector<double> pref, siteop_el;
    std::vector<QN> qnumbers;
    std::vector<std::vector<int64_t>> hss_bb;
    std::vector<const int64_t*> hss_bb_ptr;
    int physdim = 0, nr_siteops = 0, is_left = 0, nrhss = 0;
    bool built = false;
};

struct PUpdateStore {
    SymsecsFlat site_ss[3], iss, mpo;
    SiteTensor T;
    std::vector<int> bbh;
    std::vector<QN> qnumbers;
    std::vector<std::vector<int64_t>> hss_bb;
    std::vector<const int64_t*> hss_bb_ptr;
    std::vector<const int64_t*> bb_ptr;
    std::vector<const double*> tel_ptr;
    int is_left = 0, nrhss = 0, maxdim = -1;
    uint64_t seed = 0;
    bool built = false;
};

struct t3ns_ctx {
    std::unique_ptr<Context> c;
    std::vector<std::unique_ptr<BUpdateStore>> bups;
    std::vector<std::unique_ptr<QRStore>> qrs;
    std::vector<std::unique_ptr<AppendStore>> apps;
    std::vector<std::unique_ptr<PUpdateStore>> pups;
    std::vector<std::unique_ptr<SiteTensor>> sites;
};

struct MakesiteStore {
    SymsecsFlat new_ss[4][3], old_ss[4][3];
    SiteTensor sites[4];
    uint64_t seed = 0;
    bool built = false;
};

struct SplitStore {
    SymsecsFlat ss[4][3];
    std::vector<double> A;
};

struct t3ns_step {
    t3ns_ctx* ctx = nullptr;
    SplitStore spl;
    AppendStore app[3];
    PUpdateStore pup[3];
    MakesiteStore mks;
    MultisiteInfo md;
    SiteTensor site;
    ROperators ops[3];
    const Context::RawMerge* merge = nullptr;  // owned by the context
    std::vector<int> hermitian;
    std::vector<unsigned char> singlet;
    SymsecsFlat mpo;
    SymsecsFlat symarr[4][3];
    int rops_on_site[3] = {-1, -1, -1};
    int site_is_phys[4] = {0, 0, 0, 0};
    std::vector<const int64_t*> bb_ptr[3];
    std::vector<const double*> tel_ptr[3];
};

void build_append(Context& c, int pbond, int is_left, uint64_t seed, AppendStore& a) {
    // the appended site and the far bond of its operators
    const int site = c.netw.bonds[pbond][!is_left];
    int sb[3];
    c.netw.bonds_of_site(site, sb);
    const int b = sb[is_left ? 0 : 2];
    int nrops;
    std::vector<int> hss;
    ROperators o;
    if (c.netw.bonds[b][!is_left] == -1) {
        o = vacuum_rOperators(c, b, is_left);
    } else {
        std::vector<std::vector<int64_t>> tmpbb;
        operator_types(c, b, is_left, &nrops, &hss);
        o = init_rOperators_structure(c, b, is_left, false, &tmpbb);
        o.nrops = nrops;
        o.hss_of_ops = hss;
        o.beginblock.resize(nrops);
        for (int k = 0; k < nrops; ++k) o.beginblock[k] = tmpbb[hss[k]];
        o.finalize_offsets();
        o.tel.resize(o.total_size());
        Rng rng(seed);
        if (c.fill_seeded)
            for (auto& v : o.tel) v = rng.normal();
    }
    a.old = std::move(o);
    a.is_left = is_left;
    for (int k = 0; k < 3; ++k) a.site_ss[k].set(c.symsecs(sb[k]));
    a.new_ss[0].set(c.symsecs(c.get_braT3NSbond(pbond)));
    a.new_ss[1].set(c.symsecs(c.get_ketT3NSbond(pbond)));
    a.new_ss[2].set(c.symsecs(-1));
    a.old_ss[0].set(c.symsecs(c.get_braT3NSbond(b)));
    a.old_ss[1].set(c.symsecs(c.get_ketT3NSbond(b)));
    a.old_ss[2].set(c.symsecs(-1));
    a.mpo.set(c.symsecs(-1));
    const InstructionSet& ps = c.fetch_pUpdate(b, is_left);
    int nops = 0, nso = 0;
    for (const auto& i : ps.instr) {
        for (int k = 0; k < 3; ++k) a.instr.push_back(i.instr[k]);
        a.pref.push_back(i.pref);
        nops = std::max(nops, i.instr[2] + 1);
        nso = std::max(nso, i.instr[1] + 1);
    }
    a.hss_of_new.assign(ps.hss_of_new.begin(), ps.hss_of_new.begin() + nops);
    a.nr_siteops = nso;
    const Symsecs& phys = c.symsecs(sb[1]);
    a.physdim = phys.nrSecs;
    // site-operator ids may be sparse labels: tabulate the ones the
    // instructions use
    std::vector<char> used(nso, 0);
    for (const auto& i : ps.instr) used[i.instr[1]] = 1;
    a.siteop_hss.assign(nso, -1);
    a.siteop_el.assign((size_t)nso * a.physdim * a.physdim, 0.0);
    for (int so = 0; so < nso; ++so) {
        if (!used[so]) continue;
        a.siteop_hss[so] = c.ham->symsec_siteop(so, site);
        for (int br = 0; br < a.physdim; ++br)
            for (int kt = 0; kt < a.physdim; ++kt)
                a.siteop_el[((size_t)so * a.physdim + br) * a.physdim + kt] =
                    c.ham->el_siteop(so, br, kt);
    }
    const int H = c.ham->nr_hss();
    a.tprod_begin.assign(1, 0);
    for (int h = 0; h < H; ++h) {
        for (const auto& pr : c.ham->tprods_ham(h, site)) {
            a.tprods.push_back(pr[0]);
            a.tprods.push_back(pr[1]);
        }
        a.tprod_begin.push_back((int)a.tprods.size() / 2);
    }
    std::vector<std::vector<int64_t>> tmpbb;
    ROperators ns = init_rOperators_structure(c, pbond, is_left, true, &tmpbb);
    a.nrhss = ns.nrhss;
    a.bbh = ns.begin_blocks_of_hss;
    a.qnumbers = ns.qnumbers;
    a.hss_bb = tmpbb;
    for (auto& v : a.hss_bb)
        if (v.empty()) v.assign(1, 0);
    a.hss_bb_ptr.resize(a.hss_bb.size());
    for (size_t h = 0; h < a.hss_bb.size(); ++h) a.hss_bb_ptr[h] = a.hss_bb[h].data();
    a.old_bb_ptr.resize(a.old.nrops);
    a.old_tel_ptr.resize(a.old.nrops);
    for (int k = 0; k < a.old.nrops; ++k) {
        a.old_bb_ptr[k] = a.old.beginblock[k].data();
        a.old_tel_ptr[k] = a.old.tel.data() + a.old.op_offset[k];
    }
    a.built = true;
}

void fill_append(const Context& c, AppendStore& a, t3ns_append_desc* d, t3ns_rops_desc* old) {
    std::memset(d, 0, sizeof(*d));
    d->nrSyms = c.sym.nrSyms;
    for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
    d->is_left = a.is_left;
    for (int k = 0; k < 3; ++k) {
        d->site_ss[k] = a.site_ss[k].desc;
        d->new_ss[k] = a.new_ss[k].desc;
        d->old_ss[k] = a.old_ss[k].desc;
    }
    d->mpo = a.mpo.desc;
    d->nr_instr = (int)a.pref.size();
    d->instr = a.instr.data();
    d->pref = a.pref.data();
    d->hss_of_new = a.hss_of_new.data();
    d->nr_siteops = a.nr_siteops;
    d->siteop_hss = a.siteop_hss.data();
    d->physdim = a.physdim;
    d->siteop_el = a.siteop_el.data();
    d->tprod_begin = a.tprod_begin.data();
    d->tprods = a.tprods.data();
    d->nrhss = a.nrhss;
    d->begin_blocks_of_hss = a.bbh.data();
    d->qnumbers = a.qnumbers.data();
    d->hss_beginblock = a.hss_bb_ptr.data();
    std::memset(old, 0, sizeof(*old));
    old->bond = a.old.bond;
    old->is_left = a.old.is_left;
    old->P_operator = a.old.P_operator;
    old->nrhss = a.old.nrhss;
    old->begin_blocks_of_hss = a.old.begin_blocks_of_hss.data();
    old->qnumbers = a.old.qnumbers.data();
    old->nrops = a.old.nrops;
    old->hss_of_ops = a.old.hss_of_ops.data();
    old->beginblock = a.old_bb_ptr.data();
    old->tel = a.old_tel_ptr.data();
}

void build_pupdate(Context& c, int pbond, int is_left, uint64_t seed, int maxdim, PUpdateStore& u) {
    u.seed = seed;
    u.maxdim = maxdim;
    const int site = c.netw.bonds[pbond][!is_left];
    int sb[3];
    c.netw.bonds_of_site(site, sb);
    // The internal sectors the physical operators were built on and
    // a truncated set as a decomposition of the optimised tensor
    // leaves it: sectors dropped, kept with dim 0, or kept with
    // dims 1..maxdim.  maxdim < 0 keeps the context's own sectors.
    const Symsecs iss = c.symsecs(pbond);
    Symsecs tr;
    tr.bond = iss.bond;
    Rng rng(seed);
    Symsecs& slot = c.symsecs_mut(pbond);
    if (maxdim < 0) {
        tr = iss;
    } else if (maxdim == 0) {
        // no truncation: every sector keeps the full dimension of the
        // (far bond x physical) product space that couples to it
        tr = iss;
        for (auto& v : tr.dims) v = 1;
        slot = tr;
        SiteTensor unit;
        try {
            unit = init_1siteTensor(c, site);
        } catch (...) {
            slot = iss;
            throw;
        }
        const int n0 = c.symsecs(sb[0]).nrSecs, n1 = c.symsecs(sb[1]).nrSecs;
        std::fill(tr.dims.begin(), tr.dims.end(), 0);
        for (int b = 0; b < unit.nrblocks; ++b) {
            const QN q = unit.qnumbers[b];
            const int sec = is_left ? (int)(q / ((QN)n0 * n1)) : (int)(q % n0);
            tr.dims[sec] += (int)unit.block_size(b);
        }
        tr.totaldims = 0;
        for (int v : tr.dims) tr.totaldims += v;
    } else {
        for (int i = 0; i < iss.nrSecs; ++i) {
            const double r = rng.uniform();
            if (r < 0.1) continue;
            const int dim = r < 0.2 ? 0 : 1 + (int)(rng.next() % (uint64_t)maxdim);
            tr.irreps.push_back(iss.irreps[i]);
            tr.dims.push_back(dim);
            tr.fcidims.push_back(iss.fcidims.empty() ? 0.0 : iss.fcidims[i]);
            tr.totaldims += dim;
        }
        tr.nrSecs = (int)tr.dims.size();
    }
    slot = tr;
    try {
        u.T = init_1siteTensor(c, site);
        if (c.fill_seeded)
            for (auto& v : u.T.tel) v = rng.normal();
        std::vector<std::vector<int64_t>> tmpbb;
        ROperators ns = init_rOperators_structure(c, pbond, is_left, false, &tmpbb);
        u.nrhss = ns.nrhss;
        u.bbh = ns.begin_blocks_of_hss;
        u.qnumbers = ns.qnumbers;
        u.hss_bb = tmpbb;
        for (int k = 0; k < 3; ++k) u.site_ss[k].set(c.symsecs(sb[k]));
    } catch (...) {
        slot = iss;
        throw;
    }
    slot = iss;
    u.iss.set(iss);
    u.mpo.set(c.symsecs(-1));
    u.is_left = is_left;
    for (auto& v : u.hss_bb)
        if (v.empty()) v.assign(1, 0);
    u.hss_bb_ptr.resize(u.hss_bb.size());
    for (size_t h = 0; h < u.hss_bb.size(); ++h) u.hss_bb_ptr[h] = u.hss_bb[h].data();
}

void fill_pupdate(const Context& c, PUpdateStore& u, t3ns_pupdate_desc* d) {
    std::memset(d, 0, sizeof(*d));
    d->nrSyms = c.sym.nrSyms;
    for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
    d->is_left = u.is_left;
    for (int k = 0; k < 3; ++k) d->site_ss[k] = u.site_ss[k].desc;
    d->internal_ss = u.iss.desc;
    d->mpo = u.mpo.desc;
    d->T_nrblocks = u.T.nrblocks;
    d->T_qnumbers = u.T.qnumbers.data();
    d->T_beginblock = u.T.beginblock.data();
    d->T_tel = u.T.tel.data();
    d->nrhss = u.nrhss;
    d->begin_blocks_of_hss = u.bbh.data();
    d->qnumbers = u.qnumbers.data();
    d->hss_beginblock = u.hss_bb_ptr.data();
}

extern "C" {

int t3ns_ctx_create(t3ns_ctx** out, const char* netw_text, const char* model, const double* params,
                    int nparams, const int* sgs, const int* ts, int nsym, int max_dim, int minocc) {
    if (!out || !netw_text || !model) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Network n = Network::from_text(netw_text);
        std::unique_ptr<Model> m;
        const std::string mname = model;
        std::vector<int> vs(sgs, sgs + nsym), vt(ts, ts + nsym);
        bool su2 = std::find(vs.begin(), vs.end(), (int)SU2) != vs.end();
        if (mname == "NN_HUBBARD") {
            if (nparams < 2) throw std::invalid_argument("NN_HUBBARD needs t and U");
            m = make_hubbard_model(params[0], params[1], su2, nparams > 2 && params[2] != 0);
        } else {
            throw std::invalid_argument("unknown model " + mname);
        }
        auto ctx = std::make_unique<t3ns_ctx>();
        ctx->c = make_context(std::move(n), std::move(m), vs, vt);
        ctx->c->prepare_bookkeeper(max_dim, true, minocc);
        *out = ctx.release();
    });
}

int t3ns_ctx_create_qc(t3ns_ctx** out, const char* netw_text, const char* fcidump_text,
                       const int* sgs, const int* ts, int nsym, int max_dim, int minocc) {
    if (!out || !netw_text || !fcidump_text) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Network n = Network::from_text(netw_text);
        std::vector<int> vs(sgs, sgs + nsym), vt(ts, ts + nsym);
        const bool su2 = std::find(vs.begin(), vs.end(), (int)SU2) != vs.end();
        if (std::find(vs.begin(), vs.end(), (int)SENIORITY) != vs.end())
            throw std::invalid_argument("QC: seniority symmetry is not supported");
        auto ctx = std::make_unique<t3ns_ctx>();
        ctx->c = make_context(std::move(n), make_qc_model(fcidump_text, su2), vs, vt);
        ctx->c->prepare_bookkeeper(max_dim, true, minocc);
        *out = ctx.release();
    });
}

int t3ns_ctx_sweep_steps(t3ns_ctx* c, int maxsites, int* sites_out, int* nsites_out, int cap) {
    if (!c || maxsites < 1) return T3NS_ERR_INVALID;
    int n = 0;
    const int rc = guarded_host([&] {
        Network& nw = c->c->netw;
        nw.reset_sweep_state();
        StepSpecs sp;
        while (nw.next_opt_step(maxsites, &sp)) {
            if (n < cap) {
                for (int i = 0; i < STEPSPECS_MSITES; ++i)
                    sites_out[4 * n + i] = i < sp.nr_sites_opt ? sp.sites_opt[i] : -1;
                nsites_out[n] = sp.nr_sites_opt;
            }
            ++n;
        }
        nw.reset_sweep_state();
    });
    return rc != T3NS_OK ? rc : n;
}

int t3ns_ctx_destroy(t3ns_ctx* c) {
    delete c;
    return T3NS_OK;
}

int t3ns_ctx_info(const t3ns_ctx* c, int* nr_bonds, int* sites, int* psites, int* nrsyms, int* sgs) {
    if (!c) return T3NS_ERR_INVALID;
    const Context& x = *c->c;
    if (nr_bonds) *nr_bonds = x.netw.nr_bonds;
    if (sites) *sites = x.netw.sites;
    if (psites) *psites = x.netw.psites;
    if (nrsyms) *nrsyms = x.sym.nrSyms;
    if (sgs)
        for (int i = 0; i < x.sym.nrSyms; ++i) sgs[i] = x.sym.sgs[i];
    return T3NS_OK;
}

int t3ns_ctx_bond_sites(const t3ns_ctx* c, int bond, int* sites2) {
    if (!c || !sites2 || bond < 0 || bond >= c->c->netw.nr_bonds) return T3NS_ERR_INVALID;
    sites2[0] = c->c->netw.bonds[bond][0];
    sites2[1] = c->c->netw.bonds[bond][1];
    return T3NS_OK;
}

int t3ns_ctx_site_bonds(const t3ns_ctx* c, int site, int* bonds3, int* is_phys) {
    if (!c || !bonds3 || site < 0 || site >= c->c->netw.sites) return T3NS_ERR_INVALID;
    c->c->netw.bonds_of_site(site, bonds3);
    if (is_phys) *is_phys = c->c->netw.is_psite(site) ? 1 : 0;
    return T3NS_OK;
}

int t3ns_ctx_fill_seeded(t3ns_ctx* c, int on) {
    if (!c) return T3NS_ERR_INVALID;
    c->c->fill_seeded = on != 0;
    return T3NS_OK;
}

int t3ns_ctx_set_bond(t3ns_ctx* c, int bond, int nrsecs, const int* irreps, const int* dims) {
    if (!c || nrsecs < 0 || (nrsecs && (!irreps || !dims))) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& x = *c->c;
        if (bond < 0 || bond >= x.netw.nr_bonds) throw std::invalid_argument("set_bond: bad bond");
        Symsecs& slot = x.symsecs_mut(bond);
        Symsecs ns;
        ns.bond = slot.bond;
        ns.nrSecs = nrsecs;
        ns.irreps.resize(nrsecs);
        ns.dims.assign(dims, dims + nrsecs);
        ns.fcidims.assign(nrsecs, 0.0);
        for (int i = 0; i < nrsecs; ++i) {
            if (dims[i] < 0) throw std::invalid_argument("set_bond: negative dimension");
            for (int k = 0; k < T3NS_MAX_SYMMETRIES; ++k)
                ns.irreps[i][k] = irreps[i * T3NS_MAX_SYMMETRIES + k];
            // keep the full-CI dimension of a sector the bond had before
            for (int j = 0; j < slot.nrSecs; ++j)
                if (x.irreps_equal(slot.irreps[j].data(), ns.irreps[i].data()) && !slot.fcidims.empty())
                    ns.fcidims[i] = slot.fcidims[j];
            ns.totaldims += dims[i];
        }
        slot = std::move(ns);
    });
}

int t3ns_ctx_site_tensor(t3ns_ctx* c, int site, uint64_t seed, int* nrblocks, const int64_t** qnumbers,
                         const int64_t** beginblock, double** tel) {
    if (!c || !nrblocks || !qnumbers || !beginblock || !tel) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& x = *c->c;
        if (site < 0 || site >= x.netw.sites) throw std::invalid_argument("site_tensor: bad site");
        auto t = std::make_unique<SiteTensor>(init_1siteTensor(x, site));
        Rng rng(seed);
        for (auto& v : t->tel) v = rng.normal();
        *nrblocks = t->nrblocks;
        *qnumbers = t->qnumbers.data();
        *beginblock = t->beginblock.data();
        *tel = t->tel.data();
        c->sites.push_back(std::move(t));
    });
}

int t3ns_ctx_bond(const t3ns_ctx* c, int bond, int* nrsecs, int* totaldims, int* dims_out,
                  int dims_cap) {
    if (!c) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        const Symsecs& s = c->c->symsecs(bond);
        if (nrsecs) *nrsecs = s.nrSecs;
        int tot = 0;
        for (int i = 0; i < s.nrSecs; ++i) {
            tot += s.dims.empty() ? 0 : s.dims[i];
            if (dims_out && i < dims_cap) dims_out[i] = s.dims.empty() ? 0 : s.dims[i];
        }
        if (totaldims) *totaldims = tot;
    });
}

int t3ns_step_random(t3ns_step** out, t3ns_ctx* cx, const int* sites, int nsites, uint64_t seed) {
    if (!out || !cx || !sites) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *cx->c;
        auto st = std::make_unique<t3ns_step>();
        st->ctx = cx;
        MultisiteInfo& md = st->md;
        st->site = make_multisite_structure(c, sites, nsites, &md);
        SiteTensor& T = st->site;
        Rng rng(seed);
        T.tel.resize(T.size());
        // External bonds of the object, sorted (src/network.c:614-649).
        std::vector<int> ext;
        for (int i = 0; i < T.nrsites; ++i)
            for (int j = 0; j < 3; ++j) {
                const int b = md.bonds[i][j];
                if (c.netw.is_pbond(b)) continue;
                int cnt = 0;
                for (int k = 0; k < T.nrsites; ++k)
                    for (int l = 0; l < 3; ++l) cnt += md.bonds[k][l] == b;
                if (cnt == 1) ext.push_back(b);
            }
        std::sort(ext.begin(), ext.end());
        const bool isdmrg = ext.size() == 2;
        auto in_obj = [&](int s) { return std::find(T.sites, T.sites + T.nrsites, s) != T.sites + T.nrsites; };
        for (size_t i = 0; i < ext.size(); ++i) {
            const int b = ext[i];
            const int is_left = in_obj(c.netw.bonds[b][1]) ? 1 : 0;
            const int adj = c.netw.bonds[b][is_left];
            int nrops;
            std::vector<int> hss;
            ROperators r;
            std::vector<std::vector<int64_t>> tmpbb;
            if (c.netw.is_psite(adj)) {
                // Physical operator: the append of site `adj` to rops[b].
                int ab[3];
                c.netw.bonds_of_site(adj, ab);
                const int pbond = ab[is_left ? 2 : 0];
                const InstructionSet& ps = c.fetch_pUpdate(b, is_left);
                nrops = 0;
                for (const auto& ins : ps.instr) nrops = std::max(nrops, ins.instr[2] + 1);
                hss.assign(ps.hss_of_new.begin(), ps.hss_of_new.begin() + nrops);
                r = init_rOperators_structure(c, pbond, is_left, true, &tmpbb);
            } else if (c.netw.bonds[b][!is_left] == -1) {
                r = vacuum_rOperators(c, b, is_left);
                st->ops[i] = std::move(r);
                continue;
            } else {
                operator_types(c, b, is_left, &nrops, &hss);
                r = init_rOperators_structure(c, b, is_left, false, &tmpbb);
            }
            r.nrops = nrops;
            r.hss_of_ops = hss;
            r.beginblock.resize(nrops);
            for (int o = 0; o < nrops; ++o) r.beginblock[o] = tmpbb[hss[o]];
            r.finalize_offsets();
            r.tel.resize(r.total_size());
            if (c.fill_seeded)
                for (auto& v : r.tel) v = rng.normal();
            st->ops[i] = std::move(r);
        }
        // Heff descriptor pieces
        for (int i = 0; i < T.nrsites; ++i) {
            st->site_is_phys[i] = c.netw.is_psite(T.sites[i]);
            int b[3];
            c.netw.bonds_of_site(T.sites[i], b);
            for (int k = 0; k < 3; ++k) st->symarr[i][k].set(c.symsecs(b[k]));
        }
        for (int i = 0; i < (isdmrg ? 2 : 3); ++i) {
            const int s = rOperators_site_to_attach(c, st->ops[i]);
            for (int j = 0; j < T.nrsites; ++j)
                if (T.sites[j] == s) st->rops_on_site[i] = j;
        }
        st->mpo.set(c.symsecs(-1));
        const int H = c.ham->nr_hss();
        for (int h = 0; h < H; ++h) st->hermitian.push_back(c.ham->hermitian_hss(h));
        st->singlet.assign((size_t)H * H * H, 0);
        for (int a = 0; a < H; ++a)
            for (int b2 = 0; b2 < H; ++b2)
                for (int d = 0; d < H; ++d) {
                    const int m[3] = {a, b2, d};
                    st->singlet[a + (size_t)b2 * H + (size_t)d * H * H] = c.ham->mpo_couples_to_singlet(m);
                }
        st->merge = &c.fetch_raw_merge(st->ops[0].bond, isdmrg);
        if (c.fill_seeded)
            for (auto& v : T.tel) v = rng.normal();
        *out = st.release();
    });
}

int t3ns_step_append_desc(t3ns_step* st, int leg, uint64_t seed, t3ns_append_desc* d,
                          t3ns_rops_desc* old) {
    if (!st || !d || !old || leg < 0 || leg > 2) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *st->ctx->c;
        const ROperators& P = st->ops[leg];
        if (P.nrhss <= 0 || !P.P_operator)
            throw std::invalid_argument("step leg is not a physical operator");
        AppendStore& a = st->app[leg];
        if (!a.built) build_append(c, P.bond, P.is_left, seed, a);
        fill_append(c, a, d, old);
    });
}

int t3ns_step_pupdate_desc(t3ns_step* st, int leg, uint64_t seed, int maxdim,
                           t3ns_pupdate_desc* d, t3ns_rops_desc* pops) {
    if (!st || !d || !pops || leg < 0 || leg > 2 || maxdim < 0) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *st->ctx->c;
        const ROperators& P = st->ops[leg];
        if (P.nrhss <= 0 || !P.P_operator)
            throw std::invalid_argument("step leg is not a physical operator");
        PUpdateStore& u = st->pup[leg];
        if (u.built && (u.seed != seed || u.maxdim != maxdim)) u = PUpdateStore{};
        if (!u.built) {
            build_pupdate(c, P.bond, P.is_left, seed, maxdim, u);
            u.bb_ptr.resize(P.nrops);
            u.tel_ptr.resize(P.nrops);
            for (int k = 0; k < P.nrops; ++k) {
                u.bb_ptr[k] = P.beginblock[k].data();
                u.tel_ptr[k] = P.tel.data() + P.op_offset[k];
            }
            u.built = true;
        }
        fill_pupdate(c, u, d);
        std::memset(pops, 0, sizeof(*pops));
        pops->bond = P.bond;
        pops->is_left = P.is_left;
        pops->P_operator = 1;
        pops->nrhss = P.nrhss;
        pops->begin_blocks_of_hss = P.begin_blocks_of_hss.data();
        pops->qnumbers = P.qnumbers.data();
        pops->nrops = P.nrops;
        pops->hss_of_ops = P.hss_of_ops.data();
        pops->beginblock = u.bb_ptr.data();
        pops->tel = u.tel_ptr.data();
    });
}

int t3ns_step_makesite_desc(t3ns_step* st, uint64_t seed, t3ns_makesite_desc* d) {
    if (!st || !d) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *st->ctx->c;
        const SiteTensor& M = st->site;
        if (M.nrsites < 2) throw std::invalid_argument("step has a single site");
        MakesiteStore& m = st->mks;
        if (m.built && m.seed != seed) m = MakesiteStore{};
        if (!m.built) {
            m.seed = seed;
            Rng rng(seed);
            for (int i = 0; i < M.nrsites; ++i) {
                const Symsecs* ss[3] = {&st->md.ssarr_old[i][0], &st->md.ssarr_old[i][1],
                                        &st->md.ssarr_old[i][2]};
                m.sites[i] = init_1siteTensor(c, M.sites[i], ss);
                if (c.fill_seeded)
                    for (auto& v : m.sites[i].tel) v = rng.normal();
                for (int k = 0; k < 3; ++k) {
                    m.new_ss[i][k].set(st->md.ssarr[i][k]);
                    m.old_ss[i][k].set(st->md.ssarr_old[i][k]);
                }
            }
            m.built = true;
        }
        std::memset(d, 0, sizeof(*d));
        d->nrsites = M.nrsites;
        d->innersite = st->md.innersite;
        for (int i = 0; i < M.nrsites; ++i) {
            for (int k = 0; k < 3; ++k) {
                d->new_ss[i][k] = m.new_ss[i][k].desc;
                d->old_ss[i][k] = m.old_ss[i][k].desc;
            }
            d->site_nrblocks[i] = m.sites[i].nrblocks;
            d->site_qnumbers[i] = m.sites[i].qnumbers.data();
            d->site_beginblock[i] = m.sites[i].beginblock.data();
            d->site_tel[i] = m.sites[i].tel.data();
        }
        d->nrblocks = M.nrblocks;
        d->qnumbers = M.qnumbers.data();
        d->beginblock = M.beginblock.data();
    });
}

int t3ns_ctx_bupdate_desc(t3ns_ctx* cx, int site, int update_case, uint64_t seed,
                          t3ns_bupdate_desc* d, t3ns_rops_desc* ops) {
    if (!cx || !d || !ops || update_case < 0 || update_case > 2) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *cx->c;
        if (site < 0 || site >= c.netw.sites || c.netw.is_psite(site))
            throw std::invalid_argument("not a branching site");
        auto st = std::make_unique<BUpdateStore>();
        BUpdateStore& u = *st;
        u.update_case = update_case;
        int sb[3];
        c.netw.bonds_of_site(site, sb);
        // roles: ops[0] on bond (case 0 ? 1 : 0), left; ops[1] on bond
        // (case 2 ? 1 : 2), left only for case 2 (src/rOperators_bUpdate.c:292-313)
        const int obond[2] = {sb[update_case == 0 ? 1 : 0], sb[update_case == 2 ? 1 : 2]};
        const int oleft[2] = {1, update_case == 2};
        Rng rng(seed);
        for (int r = 0; r < 2; ++r) {
            ROperators o;
            if (c.netw.bonds[obond[r]][!oleft[r]] == -1) {
                o = vacuum_rOperators(c, obond[r], oleft[r]);
            } else {
                int nrops;
                std::vector<int> hss;
                std::vector<std::vector<int64_t>> tmpbb;
                operator_types(c, obond[r], oleft[r], &nrops, &hss);
                o = init_rOperators_structure(c, obond[r], oleft[r], false, &tmpbb);
                o.nrops = nrops;
                o.hss_of_ops = hss;
                o.beginblock.resize(nrops);
                for (int k = 0; k < nrops; ++k) o.beginblock[k] = tmpbb[hss[k]];
                o.finalize_offsets();
                o.tel.resize(o.total_size());
                if (c.fill_seeded)
                    for (auto& v : o.tel) v = rng.normal();
            }
            u.ops[r] = std::move(o);
            u.bb_ptr[r].resize(u.ops[r].nrops);
            u.tel_ptr[r].resize(u.ops[r].nrops);
            for (int k = 0; k < u.ops[r].nrops; ++k) {
                u.bb_ptr[r][k] = u.ops[r].beginblock[k].data();
                u.tel_ptr[r][k] = u.ops[r].tel.data() + u.ops[r].op_offset[k];
            }
        }
        u.T = init_1siteTensor(c, site);
        if (c.fill_seeded)
            for (auto& v : u.T.tel) v = rng.normal();
        for (int k = 0; k < 3; ++k) u.site_ss[k].set(c.symsecs(sb[k]));
        u.mpo.set(c.symsecs(-1));
        const int nbond = sb[update_case];
        const int nleft = update_case == 2;
        const InstructionSet& bs = c.fetch_bUpdate(nbond, nleft);
        int nops = 0;
        for (const auto& i : bs.instr) {
            for (int k = 0; k < 3; ++k) u.instr.push_back(i.instr[k]);
            u.pref.push_back(i.pref);
            nops = std::max(nops, i.instr[2] + 1);
        }
        u.hss_of_new.assign(bs.hss_of_new.begin(), bs.hss_of_new.begin() + nops);
        const int H = c.ham->nr_hss();
        u.tprod_begin.assign(1, 0);
        for (int h = 0; h < H; ++h) {
            for (const auto& pr : c.ham->tprods_ham(h, site)) {
                u.tprods.push_back(pr[0]);
                u.tprods.push_back(pr[1]);
            }
            u.tprod_begin.push_back((int)u.tprods.size() / 2);
        }
        std::vector<std::vector<int64_t>> tmpbb;
        ROperators ns = init_rOperators_structure(c, nbond, nleft, false, &tmpbb);
        u.nrhss = ns.nrhss;
        u.bbh = ns.begin_blocks_of_hss;
        u.qnumbers = ns.qnumbers;
        u.hss_bb = tmpbb;
        for (auto& v : u.hss_bb)
            if (v.empty()) v.assign(1, 0);
        u.hss_bb_ptr.resize(u.hss_bb.size());
        for (size_t h = 0; h < u.hss_bb.size(); ++h) u.hss_bb_ptr[h] = u.hss_bb[h].data();

        std::memset(d, 0, sizeof(*d));
        d->nrSyms = c.sym.nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
        d->update_case = update_case;
        for (int k = 0; k < 3; ++k) d->site_ss[k] = u.site_ss[k].desc;
        d->mpo = u.mpo.desc;
        d->T_nrblocks = u.T.nrblocks;
        d->T_qnumbers = u.T.qnumbers.data();
        d->T_beginblock = u.T.beginblock.data();
        d->T_tel = u.T.tel.data();
        d->nr_instr = (int)u.pref.size();
        d->instr = u.instr.data();
        d->pref = u.pref.data();
        d->hss_of_new = u.hss_of_new.data();
        d->tprod_begin = u.tprod_begin.data();
        d->tprods = u.tprods.data();
        d->nrhss = u.nrhss;
        d->begin_blocks_of_hss = u.bbh.data();
        d->qnumbers = u.qnumbers.data();
        d->hss_beginblock = u.hss_bb_ptr.data();
        for (int r = 0; r < 2; ++r) {
            t3ns_rops_desc& o = ops[r];
            std::memset(&o, 0, sizeof(o));
            o.bond = u.ops[r].bond;
            o.is_left = u.ops[r].is_left;
            o.P_operator = 0;
            o.nrhss = u.ops[r].nrhss;
            o.begin_blocks_of_hss = u.ops[r].begin_blocks_of_hss.data();
            o.qnumbers = u.ops[r].qnumbers.data();
            o.nrops = u.ops[r].nrops;
            o.hss_of_ops = u.ops[r].hss_of_ops.data();
            o.beginblock = u.bb_ptr[r].data();
            o.tel = u.tel_ptr[r].data();
        }
        cx->bups.push_back(std::move(st));
    });
}

int t3ns_ctx_append_desc(t3ns_ctx* cx, int bond, int is_left, uint64_t seed, t3ns_append_desc* d,
                         t3ns_rops_desc* old) {
    if (!cx || !d || !old) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *cx->c;
        if (bond < 0 || bond >= c.netw.nr_bonds || !c.netw.is_psite(c.netw.bonds[bond][!is_left]))
            throw std::invalid_argument("append: the bond's site on that side is not physical");
        auto a = std::make_unique<AppendStore>();
        // The physical operators live on the bond's sectors at unit dims
        // (src/optimize_network.c:440-453); the context keeps its dims.
        Symsecs& slot = c.symsecs_mut(bond);
        const Symsecs keep = slot;
        std::fill(slot.dims.begin(), slot.dims.end(), 1);
        slot.totaldims = slot.nrSecs;
        try {
            build_append(c, bond, is_left ? 1 : 0, seed, *a);
        } catch (...) {
            slot = keep;
            throw;
        }
        slot = keep;
        fill_append(c, *a, d, old);
        cx->apps.push_back(std::move(a));
    });
}

int t3ns_ctx_pupdate_desc(t3ns_ctx* cx, int bond, int is_left, uint64_t seed, t3ns_pupdate_desc* d) {
    if (!cx || !d) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *cx->c;
        if (bond < 0 || bond >= c.netw.nr_bonds || !c.netw.is_psite(c.netw.bonds[bond][!is_left]))
            throw std::invalid_argument("pupdate: the bond's site on that side is not physical");
        auto u = std::make_unique<PUpdateStore>();
        build_pupdate(c, bond, is_left ? 1 : 0, seed, -1, *u);
        u->built = true;
        fill_pupdate(c, *u, d);
        cx->pups.push_back(std::move(u));
    });
}

int t3ns_step_split_desc(t3ns_step* st, int siteV, int minD, int maxD, double truncerr,
                         char truncType, t3ns_split_desc* d) {
    if (!st || !d) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *st->ctx->c;
        const SiteTensor& M = st->site;
        if (M.nrsites < 2 || siteV < 0 || siteV >= M.nrsites)
            throw std::invalid_argument("split: bad site");
        SplitStore& sp = st->spl;
        if (sp.A.empty()) {
            sp.A = M.tel;
            double n = 0;
            for (double v : sp.A) n += v * v;
            n = std::sqrt(n);
            for (auto& v : sp.A) v /= n;
            for (int i = 0; i < M.nrsites; ++i)
                for (int k = 0; k < 3; ++k) sp.ss[i][k].set(st->md.ssarr[i][k]);
        }
        std::memset(d, 0, sizeof(*d));
        d->nrSyms = c.sym.nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
        d->nrsites = M.nrsites;
        d->siteV = siteV;
        for (int i = 0; i < M.nrsites; ++i)
            for (int k = 0; k < 3; ++k) d->ss[i][k] = sp.ss[i][k].desc;
        d->nrblocks = M.nrblocks;
        d->qnumbers = M.qnumbers.data();
        d->beginblock = M.beginblock.data();
        d->tel = sp.A.data();
        d->minD = minD;
        d->maxD = maxD;
        d->truncerr = truncerr;
        d->truncType = truncType;
    });
}

int t3ns_ctx_qr_desc(t3ns_ctx* cx, int siteA, int siteC, uint64_t seed, t3ns_qr_desc* d) {
    if (!cx || !d) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        Context& c = *cx->c;
        if (siteA < 0 || siteA >= c.netw.sites || siteC < 0 || siteC >= c.netw.sites)
            throw std::invalid_argument("qr: bad site");
        auto st = std::make_unique<QRStore>();
        QRStore& q = *st;
        Rng rng(seed);
        q.A = init_1siteTensor(c, siteA);
        for (auto& v : q.A.tel) v = rng.normal();
        q.C = init_1siteTensor(c, siteC);
        for (auto& v : q.C.tel) v = rng.normal();
        int ba[3], bc[3];
        c.netw.bonds_of_site(siteA, ba);
        c.netw.bonds_of_site(siteC, bc);
        for (int k = 0; k < 3; ++k) {
            q.ssA[k].set(c.symsecs(ba[k]));
            q.ssC[k].set(c.symsecs(bc[k]));
        }
        std::memset(d, 0, sizeof(*d));
        d->nrSyms = c.sym.nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
        for (int k = 0; k < 3; ++k) {
            d->ssA[k] = q.ssA[k].desc;
            d->ssC[k] = q.ssC[k].desc;
        }
        d->A_nrblocks = q.A.nrblocks;
        d->A_qnumbers = q.A.qnumbers.data();
        d->A_beginblock = q.A.beginblock.data();
        d->A_tel = q.A.tel.data();
        d->C_nrblocks = q.C.nrblocks;
        d->C_qnumbers = q.C.qnumbers.data();
        d->C_beginblock = q.C.beginblock.data();
        d->C_tel = q.C.tel.data();
        cx->qrs.push_back(std::move(st));
    });
}

int t3ns_step_destroy(t3ns_step* s) {
    delete s;
    return T3NS_OK;
}

int64_t t3ns_step_size(const t3ns_step* s) { return s ? s->site.size() : -1; }

int t3ns_step_heff_desc(t3ns_step* s, t3ns_heff_desc* d, t3ns_rops_desc* ops, t3ns_site_desc* site) {
    if (!s || !d || !ops || !site) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        const Context& c = *s->ctx->c;
        std::memset(d, 0, sizeof(*d));
        d->nrSyms = c.sym.nrSyms;
        for (int i = 0; i < T3NS_MAX_SYMMETRIES; ++i) d->sgs[i] = c.sym.sgs[i];
        d->mpo = s->mpo.desc;
        d->hermitian = s->hermitian.data();
        d->singlet = s->singlet.data();
        const SiteTensor& T = s->site;
        for (int i = 0; i < T.nrsites; ++i)
            for (int k = 0; k < 3; ++k) d->symarr[i][k] = s->symarr[i][k].desc;
        for (int k = 0; k < 3; ++k) d->rops_on_site[k] = s->rops_on_site[k];
        d->merge_step = s->merge->set.step;
        d->merge_nr = s->merge->set.nr_instr();
        d->merge_instr = s->merge->instr.data();
        d->merge_pref = s->merge->pref.data();
        for (int k = 0; k < 3; ++k) {
            const ROperators& r = s->ops[k];
            t3ns_rops_desc& o = ops[k];
            std::memset(&o, 0, sizeof(o));
            if (r.nrhss <= 0) continue;
            o.bond = r.bond;
            o.is_left = r.is_left;
            o.P_operator = r.P_operator;
            o.nrhss = r.nrhss;
            o.begin_blocks_of_hss = r.begin_blocks_of_hss.data();
            o.qnumbers = r.qnumbers.data();
            o.nrops = r.nrops;
            o.hss_of_ops = r.hss_of_ops.data();
            s->bb_ptr[k].resize(r.nrops);
            s->tel_ptr[k].resize(r.nrops);
            for (int op = 0; op < r.nrops; ++op) {
                s->bb_ptr[k][op] = r.beginblock[op].data();
                s->tel_ptr[k][op] = r.tel.data() + r.op_offset[op];
            }
            o.beginblock = s->bb_ptr[k].data();
            o.tel = s->tel_ptr[k].data();
        }
        std::memset(site, 0, sizeof(*site));
        site->nrsites = T.nrsites;
        for (int i = 0; i < 4; ++i) {
            site->sites[i] = T.sites[i];
            site->site_is_phys[i] = s->site_is_phys[i];
        }
        site->nrblocks = T.nrblocks;
        site->qnumbers = T.qnumbers.data();
        site->beginblock = T.beginblock.data();
    });
}

int t3ns_ckpt_write(const t3ns_ctx* c, const char* path, int nsites, const int* nrblocks,
                    const int64_t* const* qnumbers, const int64_t* const* beginblock,
                    const double* const* tel) {
    if (!c || !path || nsites < 0 || (nsites > 0 && (!nrblocks || !qnumbers || !beginblock || !tel)))
        return T3NS_ERR_INVALID;
    return guarded_host([&] {
        if (nsites != c->c->netw.sites)
            throw std::invalid_argument("checkpoint: one tensor per network site expected");
        std::vector<SiteBlocks> ts(nsites);
        for (int i = 0; i < nsites; ++i) {
            const int nb = nrblocks[i];
            if (nb < 0 || (nb > 0 && (!qnumbers[i] || !beginblock[i])))
                throw std::invalid_argument("checkpoint: bad site tensor " + std::to_string(i));
            SiteBlocks& t = ts[i];
            t.sites = {i};
            t.qnumbers.assign(qnumbers[i], qnumbers[i] + nb);
            t.beginblock.assign(1, 0);
            if (nb > 0) t.beginblock.assign(beginblock[i], beginblock[i] + nb + 1);
            const int64_t n = t.beginblock.back();
            if (n < 0) throw std::invalid_argument("checkpoint: bad block offsets");
            if (n > 0 && !tel[i]) throw std::invalid_argument("checkpoint: missing elements");
            t.tel.assign(tel[i], tel[i] + n);
        }
        write_checkpoint(path, *c->c, ts);
    });
}

struct t3ns_ckpt {
    CheckpointData d;
};

int t3ns_ckpt_open(t3ns_ckpt** k, const char* path) {
    if (!k || !path) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        auto p = std::make_unique<t3ns_ckpt>();
        p->d = read_checkpoint(path);
        *k = p.release();
    });
}

int t3ns_ckpt_info(const t3ns_ckpt* k, int* sites, int* nr_bonds, int* ham_type,
                   int* nr_rops_groups) {
    if (!k) return T3NS_ERR_INVALID;
    if (sites) *sites = k->d.netw.sites;
    if (nr_bonds) *nr_bonds = k->d.netw.nr_bonds;
    if (ham_type) *ham_type = k->d.ham.type;
    if (nr_rops_groups) *nr_rops_groups = k->d.nr_rops_groups;
    return T3NS_OK;
}

int t3ns_ckpt_site(const t3ns_ckpt* k, int site, int* nrsites, int* nrblocks,
                   const int64_t** qnumbers, const int64_t** beginblock, const double** tel) {
    if (!k || site < 0 || site >= (int)k->d.tensors.size()) return T3NS_ERR_INVALID;
    const SiteBlocks& t = k->d.tensors[site];
    if (nrsites) *nrsites = (int)t.sites.size();
    if (nrblocks) *nrblocks = t.nrblocks();
    if (qnumbers) *qnumbers = t.qnumbers.data();
    if (beginblock) *beginblock = t.beginblock.data();
    if (tel) *tel = t.tel.data();
    return T3NS_OK;
}

int t3ns_ckpt_context(const t3ns_ckpt* k, t3ns_ctx** out) {
    if (!k || !out) return T3NS_ERR_INVALID;
    return guarded_host([&] {
        auto ctx = std::make_unique<t3ns_ctx>();
        ctx->c = context_from_checkpoint(k->d);
        *out = ctx.release();
    });
}

int t3ns_ckpt_close(t3ns_ckpt* k) {
    delete k;
    return T3NS_OK;
}

}  // extern "C"
