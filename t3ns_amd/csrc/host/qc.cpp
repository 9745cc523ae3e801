// Quantum-chemistry Hamiltonian: FCIDUMP integrals, the complementary-
// operator bookkeeping (operator "types" with creator/position/spin tags per
// bond) and the instruction lists for renormalisation and Heff merging.
// Follows src/hamiltonian_qc.c, src/opType.c, src/opType_qc.c and
// src/instructions_qc.c.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>

#include "model.hpp"

namespace t3 {
namespace {

constexpr int NR_OPS = 5, NR_TYP = 2, BT = 3;  // operator counts, types, tag size
// Tags per operator of each (count, type) class (src/opType_qc.c:72-81).
const int kNrBaseTags[NR_OPS][NR_TYP] = {{0, -1}, {1, -1}, {2, 2}, {3, 1}, {4, 0}};
// Which creator patterns make each class (src/opType_qc.c:56-67).
const int kTodo[NR_OPS][NR_TYP] = {{0, 0}, {2, 0}, {3, 3}, {0, 2}, {0, 0}};
const int kCreator[NR_OPS][NR_TYP][3][2] = {{{{0}}, {{0}}},
                                             {{{1}, {0}}, {{0}}},
                                             {{{1, 1}, {0, 0}, {1, 0}}, {{1, 1}, {0, 0}, {1, 0}}},
                                             {{{0}}, {{1}, {0}}},
                                             {{{0}}, {{0}}}};
// Operator-count combinations of the three legs (src/opType_qc.c:69-76).
const int kCombine[15][3] = {{4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {3, 0, 1},
                             {0, 3, 1}, {1, 3, 0}, {1, 0, 3}, {0, 1, 3}, {2, 2, 0},
                             {0, 2, 2}, {2, 0, 2}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};
// MPO-bond irreps (src/hamiltonian_qc.c:39-49).
const int kIrrU1[13][2] = {{-2, 0}, {-1, -1}, {-1, 0}, {-1, 1}, {0, -2}, {0, -1}, {0, 0},
                           {0, 1},  {0, 2},   {1, -1}, {1, 0},  {1, 1},  {2, 0}};
const int kValidU1[13] = {0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0};
const int kIrrSU2[8][2] = {{-2, 2}, {-2, 0}, {-1, 1}, {0, 2}, {0, 0}, {1, 1}, {2, 0}, {2, 2}};
const int kValidSU2[8] = {0, 1, 1, 1, 1, 1, 1, 0};
// One-site operator tags, position filled in per site (src/hamiltonian_qc.c:1287-1312,1402-1421).
const int kSiteNrU1[5] = {1, 4, 6, 4, 1};
const int kSiteTagsU1[32][3] = {
    {1, -1, 0}, {1, -1, 1}, {0, -1, 0}, {0, -1, 1},                          // 1 op
    {1, -1, 0}, {1, -1, 1}, {0, -1, 0}, {0, -1, 1}, {1, -1, 0}, {0, -1, 0},  // 2 ops
    {1, -1, 1}, {0, -1, 0}, {1, -1, 0}, {0, -1, 1}, {1, -1, 1}, {0, -1, 1},
    {1, -1, 0}, {1, -1, 1}, {0, -1, 0}, {1, -1, 0}, {1, -1, 1}, {0, -1, 1},  // 3 ops
    {1, -1, 0}, {0, -1, 0}, {0, -1, 1}, {1, -1, 1}, {0, -1, 0}, {0, -1, 1},
    {1, -1, 0}, {1, -1, 1}, {0, -1, 0}, {0, -1, 1}};                         // 4 ops
const int kSiteNrSU2[5] = {1, 2, 4, 2, 1};
const int kSiteTagsSU2[20][3] = {
    {1, -1, 1},  {0, -1, 1},                                                   // 1 op
    {1, -1, -1}, {1, -1, 0}, {0, -1, -1}, {0, -1, 0}, {1, -1, -1}, {0, -1, 0},  // 2 ops
    {1, -1, -1}, {0, -1, 2},
    {1, -1, -1}, {1, -1, -1}, {0, -1, 1}, {1, -1, -1}, {0, -1, -1}, {0, -1, 1},  // 3 ops
    {1, -1, -1}, {1, -1, -1}, {0, -1, -1}, {0, -1, 0}};                          // 4 ops

struct OpType {
    int begin[NR_OPS * NR_TYP + 1] = {0};
    std::vector<int> tags[NR_OPS][NR_TYP];
    int amount(int nrop, int typ) const { return begin[nrop * NR_TYP + typ + 1] - begin[nrop * NR_TYP + typ]; }
    void range(int nrop, int* off, int* amt) const {
        *off = begin[nrop * NR_TYP];
        *amt = std::max(0, begin[(nrop + 1) * NR_TYP] - *off);
    }
    void type_of(int id, int* nr, int* typ, int* k) const {
        for (*nr = 0; *nr < NR_OPS; ++*nr) {
            for (*typ = 0; *typ < NR_TYP; ++*typ)
                if (begin[*nr * NR_TYP + *typ + 1] > id) break;
            if (*typ != NR_TYP) break;
        }
        *k = id - begin[*nr * NR_TYP + *typ];
    }
    const int* tag(int nr, int typ, int k) const {
        const int n = kNrBaseTags[nr][typ];
        if (n <= 0) return nullptr;
        return tags[nr][typ].data() + (size_t)k * n * BT;
    }
};

class QC final : public Model {
public:
    QC(const std::string& text, bool su2) : su2_(su2) { read_fcidump(text); }
    // From checkpointed integrals (one-body part already folded in,
    // orbital irreps in psi4 order).
    explicit QC(const HamData& d) : su2_(d.su2 != 0), restored_(true) {
        norb_ = d.norb;
        orbirrep_ = d.orbirrep;
        V_ = d.Vijkl;
        core_ = d.core_energy;
        if (norb_ <= 0 || (int)orbirrep_.size() != norb_ ||
            V_.size() != (size_t)norb_ * norb_ * norb_ * norb_)
            throw std::runtime_error("QC: checkpointed integrals of the wrong size");
        if (d.has_seniority) throw std::runtime_error("QC: seniority symmetry is not supported");
    }
    std::string name() const override { return "QC"; }
    HamData ham_data() const override {
        HamData d;
        d.type = HAM_QC;
        d.su2 = su2_;
        d.norb = norb_;
        d.orbirrep = orbirrep_;
        d.Vijkl = V_;
        d.core_energy = core_;
        d.has_seniority = 0;
        return d;
    }

    void attach(const Network& n, const SymInfo& s) override {
        netw_ = &n;
        sym_ = s;
        pg_ = -1;
        for (int i = 0; i < s.nrSyms; ++i)
            if (s.sgs[i] >= C1 && s.sgs[i] < SENIORITY) pg_ = s.sgs[i];
        npg_ = pg_ == -1 ? 1 : pg_nr_irreps(pg_);
        if (norb_ != n.psites)
            throw std::runtime_error("FCIDUMP orbital count differs from the network's physical sites");
        if (restored_) {  // integrals and irreps as checkpointed
            prepare_mpo();
            init_optype_array();
            return;
        }
        orbirrep_.assign(norb_, 0);
        if (pg_ != -1) {
            if ((int)orbsym_.size() != norb_) throw std::runtime_error("FCIDUMP lacks ORBSYM");
            for (int i = 0; i < norb_; ++i) {
                if (orbsym_[i] < 1 || orbsym_[i] > npg_) throw std::runtime_error("bad ORBSYM");
                orbirrep_[i] = fcidump_to_psi4(orbsym_[i] - 1, pg_);
            }
        }
        int N = 0;
        for (int i = 0; i < s.nrSyms; ++i)
            if (s.sgs[i] == U1) N += s.target_state[i];
        form_integrals(N);
        prepare_mpo();
        init_optype_array();
    }

    Symsecs phys_symsecs(int orb) const override {
        static const int u1[4][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
        static const int s2[3][3] = {{0, 0, 0}, {1, 1, 1}, {0, 2, 0}};
        Symsecs r;
        r.nrSecs = su2_ ? 3 : 4;
        r.totaldims = r.nrSecs;
        r.irreps.assign(r.nrSecs, Irrep{});
        r.dims.assign(r.nrSecs, 1);
        r.fcidims.assign(r.nrSecs, 1.0);
        for (int i = 0; i < r.nrSecs; ++i) {
            for (int j = 0; j < 3; ++j) r.irreps[i][j] = su2_ ? s2[i][j] : u1[i][j];
            if (pg_ != -1) r.irreps[i][3] = r.irreps[i][0] ? orbirrep_[orb] : 0;
        }
        return r;
    }
    const Symsecs& mpo_symsecs() const override { return mpo_; }
    int trivial_hss() const override {
        const int n = su2_ ? 8 : 13;
        for (int i = 0; i < n; ++i)
            if (irr(i)[0] == 0 && irr(i)[1] == 0) return npg_ * i;
        return -1;
    }
    int hermitian_hss(int h) const override {
        const int pgi = h % npg_, o = h / npg_;
        const int a = -irr(o)[0], b = irr(o)[1] * (su2_ ? 1 : -1);
        const int n = su2_ ? 8 : 13;
        for (int i = 0; i < n; ++i)
            if (irr(i)[0] == a && irr(i)[1] == b) return i * npg_ + pgi;
        return -1;
    }
    bool mpo_couples_to_singlet(const int m[3]) const override {
        // src/hamiltonian_qc.c:367-383
        int p[3], o[3];
        for (int i = 0; i < 3; ++i) {
            p[i] = m[i] % npg_;
            o[i] = m[i] / npg_;
        }
        if ((p[0] ^ p[1]) != p[2]) return false;
        return valid_tprod(o[0], o[1], o[2], false);
    }
    int symsec_siteop(int siteop, int site) const override {
        OpType so = site_optype(netw_->sitetoorb[site]);
        int nr, typ, k;
        so.type_of(siteop, &nr, &typ, &k);
        return symsec_of(so, nr, typ, k);
    }
    double el_siteop(int op, int bra, int ket) const override {
        return su2_ ? su2_el(op, bra, ket) : u1_el(op, bra, ket);
    }
    std::vector<std::array<int, 2>> tprods_ham(int res, int site) const override {
        // src/hamiltonian_qc.c:336-365, 1183-1209
        const int n = su2_ ? 8 : 13;
        const int pgr = res % npg_, other = res / npg_;
        const bool ps = netw_->is_psite(site);
        std::vector<std::array<int, 2>> out;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                if (!valid_tprod(i, j, other, ps)) continue;
                if (ps) {
                    const int p1 = orbirrep_[netw_->sitetoorb[site]] * (double_op(i) == 0);
                    out.push_back({i * npg_ + p1, j * npg_ + (p1 ^ pgr)});
                } else {
                    for (int p1 = 0; p1 < npg_; ++p1) out.push_back({i * npg_ + p1, j * npg_ + (p1 ^ pgr)});
                }
            }
        return out;
    }
    bool consistent_state(const int* ts, const SymInfo& s) const override {
        // src/hamiltonian_qc.c:650-686
        int parity = 0, particles = 0, su2v = -1;
        for (int i = 0; i < s.nrSyms; ++i) {
            switch (s.sgs[i]) {
                case Z2: parity = ts[i]; break;
                case U1:
                    particles += ts[i];
                    if (ts[i] > s.target_state[i]) return false;
                    break;
                case SU2: su2v = ts[i]; break;
                default: break;
            }
        }
        if (parity % 2 != particles % 2) return false;
        if (su2v != -1 && particles % 2 != su2v % 2) return false;
        return true;
    }

    InstructionSet raw_pUpdate(const Network& n, int bond, int is_left) override {
        return fetch_pupdate(n, bond, is_left, true);
    }
    InstructionSet raw_bUpdate(const Network& n, int bond, int is_left) override {
        return fetch_bupdate(n, bond, is_left, true);
    }
    InstructionSet raw_merge(const Network& n, int bond, int isdmrg) override {
        // src/instructions_qc.c:284-326
        InstructionSet s;
        OpType ops[3];
        int order[3] = {0, 1, 2};
        if (isdmrg) {
            ops[0] = optype(bond, 1, true);
            ops[1] = unity_optype();
            ops[2] = optype(bond, 0, true);
            order[1] = 2;
            order[2] = 1;
        } else {
            const int bsite = n.bonds[bond][n.is_psite(n.bonds[bond][0])];
            int b[3];
            n.bonds_of_site(bsite, b);
            ops[0] = optype(b[0], 1, true);
            ops[1] = optype(b[1], 1, true);
            ops[2] = optype(b[2], 0, true);
        }
        s.step = 2 + !isdmrg;
        combine_all(ops, isdmrg ? 'd' : 't', order, &s);
        return s;
    }

private:
    bool su2_;
    bool restored_ = false;
    int norb_ = 0;
    std::vector<int> orbsym_, orbirrep_;
    double core_ = 0;
    std::vector<double> h1_, V_;
    const Network* netw_ = nullptr;
    SymInfo sym_;
    int pg_ = -1, npg_ = 1;
    Symsecs mpo_;
    std::map<std::pair<int, int>, OpType> filtered_;  // opType_arr

    const int* irr(int i) const { return su2_ ? kIrrSU2[i] : kIrrU1[i]; }
    bool valid_tprod(int i, int j, int r, bool psite) const {
        // src/hamiltonian_qc.c:1126-1147
        if (psite && !(su2_ ? kValidSU2[i] : kValidU1[i])) return false;
        const int* a = irr(i);
        const int* b = irr(j);
        const int* c = irr(r);
        if (su2_)
            return a[0] + b[0] == c[0] && (a[1] + b[1] + c[1]) % 2 == 0 &&
                   c[1] >= std::abs(a[1] - b[1]) && c[1] <= a[1] + b[1];
        return a[0] + b[0] == c[0] && a[1] + b[1] == c[1];
    }
    int double_op(int i) const {
        return su2_ ? std::abs(kIrrSU2[i][0]) % 2 == 0 : std::abs(kIrrU1[i][0] + kIrrU1[i][1]) % 2 == 0;
    }
    void prepare_mpo() {
        // src/hamiltonian_qc.c:1157-1181
        const int n = su2_ ? 8 : 13;
        mpo_ = Symsecs();
        mpo_.bond = -1;
        mpo_.nrSecs = n * npg_;
        mpo_.irreps.assign(mpo_.nrSecs, Irrep{});
        mpo_.dims.assign(mpo_.nrSecs, 1);
        mpo_.fcidims.assign(mpo_.nrSecs, 1.0);
        mpo_.totaldims = mpo_.nrSecs;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < npg_; ++j) {
                Irrep& r = mpo_.irreps[i * npg_ + j];
                r[0] = su2_ ? std::abs(irr(i)[0]) % 2 : std::abs(irr(i)[0] + irr(i)[1]) % 2;
                r[1] = irr(i)[0];
                r[2] = irr(i)[1];
                if (pg_ != -1) r[3] = j;
            }
    }

    // ------------------------------------------------------------ FCIDUMP
    void read_fcidump(const std::string& text) {
        // Header up to &END / /END / / (src/hamiltonian_qc.c:751-828).
        size_t pos = 0;
        std::string header;
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            std::string t;
            for (char ch : line)
                if (!isspace((unsigned char)ch)) t += ch;
            std::string up = t;
            for (auto& ch : up) ch = (char)toupper((unsigned char)ch);
            if (up == "&END" || up == "/END" || up == "/") break;
            header += line + "\n";
        }
        (void)pos;
        std::string up = header;
        for (auto& ch : up) ch = (char)toupper((unsigned char)ch);
        auto value_after = [&](const std::string& key) -> std::string {
            size_t k = up.find(key);
            if (k == std::string::npos) return "";
            size_t e = up.find('=', k);
            if (e == std::string::npos) return "";
            size_t st = e + 1;
            size_t stop = st;
            // up to the next KEY= (letters followed by '=') or end
            while (stop < up.size()) {
                if (isalpha((unsigned char)up[stop])) {
                    size_t q = stop;
                    while (q < up.size() && (isalnum((unsigned char)up[q]) || up[q] == '_')) ++q;
                    size_t r = q;
                    while (r < up.size() && isspace((unsigned char)up[r])) ++r;
                    if (r < up.size() && up[r] == '=') break;
                    stop = q;
                    continue;
                }
                if (up[stop] == '&' || up[stop] == '/') break;
                ++stop;
            }
            return up.substr(st, stop - st);
        };
        norb_ = std::atoi(value_after("NORB").c_str());
        if (norb_ <= 0) throw std::runtime_error("FCIDUMP: cannot read NORB");
        {
            std::string os = value_after("ORBSYM");
            for (auto& ch : os)
                if (ch == ',') ch = ' ';
            std::istringstream ss(os);
            int v;
            while (ss >> v) orbsym_.push_back(v);
        }
        const int n = norb_;
        h1_.assign((size_t)n * n, 0.0);
        V_.assign((size_t)n * n * n * n, 0.0);
        core_ = 0;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            double v;
            int i, j, k, l;
            if (!(ls >> v >> i >> j >> k >> l)) continue;
            if (k != 0)
                V_[(size_t)(l - 1) * n * n * n + (size_t)(k - 1) * n * n + (size_t)(j - 1) * n + (i - 1)] = v;
            else if (i != 0)
                h1_[(size_t)(j - 1) * n + (i - 1)] = v;
            else
                core_ = v;
        }
    }

    void form_integrals(int N) {
        // Eight-fold fill and one-body folding (src/hamiltonian_qc.c:903-953).
        const int n = norb_;
        const size_t n2 = (size_t)n * n, n3 = n2 * n;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) h1_[(size_t)i * n + j] = h1_[(size_t)j * n + i];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = 0; k <= i; ++k)
                    for (int l = 0; l <= k; ++l) {
                        const double v = V_[i + n * j + n2 * k + n3 * l];
                        if (std::fabs(v) < 1e-20) continue;
                        V_[k + n * l + n2 * i + n3 * j] = v;
                        V_[j + n * i + n2 * l + n3 * k] = v;
                        V_[l + n * k + n2 * j + n3 * i] = v;
                        V_[j + n * i + n2 * k + n3 * l] = v;
                        V_[l + n * k + n2 * i + n3 * j] = v;
                        V_[i + n * j + n2 * l + n3 * k] = v;
                        V_[k + n * l + n2 * j + n3 * i] = v;
                    }
        const double pref = 1.0 / (N * 1. - 1);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double p2 = pref * h1_[(size_t)i * n + j];
                for (int k = 0; k < n; ++k) {
                    V_[i + n * j + n2 * k + n3 * k] += p2;
                    V_[k + n * k + n2 * i + n3 * j] += p2;
                }
            }
    }

    double get_V(const int* t1, const int* t2, const int* t3, const int* t4) const {
        // src/hamiltonian_qc.c:578-595
        if (t1[0] != 1 || t2[0] != 1 || t3[0] != 0 || t4[0] != 0) return 0;
        if (!su2_ && (t1[2] != t4[2] || t2[2] != t3[2])) return 0;
        if (pg_ != -1 && ((orbirrep_[t1[1]] ^ orbirrep_[t2[1]]) ^ (orbirrep_[t3[1]] ^ orbirrep_[t4[1]])) != 0)
            return 0;
        const size_t n = norb_;
        return V_[t1[1] + n * t4[1] + n * n * t2[1] + n * n * n * t3[1]];
    }
    double Bv(const int* const t[4], int twoJ) const {
        const double r = -std::sqrt(twoJ + 1.0);
        return r * (get_V(t[0], t[1], t[3], t[2]) + (twoJ == 2 ? -1 : 1) * get_V(t[0], t[1], t[2], t[3]));
    }
    double Bt(const int* const t[4], int twoJ) const {
        if (twoJ == 0) return 2 * get_V(t[0], t[1], t[3], t[2]) - get_V(t[0], t[1], t[2], t[3]);
        return std::sqrt(3.0) * get_V(t[0], t[1], t[2], t[3]);
    }

    // ------------------------------------------------------- tag algebra
    int equaltags(const int* a, const int* b, int n1, int n2) const {
        // src/hamiltonian_qc.c:712-749
        const int maxj = BT - su2_;
        if (n2 == 0) return 0;
        if (n1 == 2 && n2 == 1) {
            for (int i = 0; i < n1; ++i) {
                int j;
                for (j = 0; j < maxj; ++j)
                    if (a[i * BT + j] != b[j]) break;
                if (j == maxj) return i + 1;
            }
            return 0;
        }
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < maxj; ++j)
                if (a[i * BT + j] != b[i * BT + j]) return 0;
        if (su2_) return a[n1 * BT - 1] == b[n2 * BT - 1];
        return 1;
    }
    bool compare_tags(const int* const tags[3], const int nt[3], int sumleg, double* val, bool dmrgmerge) const {
        // src/hamiltonian_qc.c:428-495
        const int ns[2] = {sumleg == 0, 1 + (sumleg != 2)};
        if (nt[sumleg] == 0) {
            *val = 1;
            return true;
        }
        if (nt[sumleg] == 1) {
            const int other = ns[nt[ns[0]] == 0];
            if (!equaltags(tags[sumleg], tags[other], nt[sumleg], nt[other])) return false;
            if (su2_) *val = dmrgmerge ? 1 : (nt[1] == 0 ? -1 : 1);
            else *val = (sumleg == 1 && other == 0) ? -1 : 1;
            return true;
        }
        if (nt[sumleg] == 2) {
            int p0 = equaltags(tags[sumleg], tags[ns[0]], nt[sumleg], nt[ns[0]]);
            if (p0 == 0 && nt[ns[0]] != 0) return false;
            int p1 = equaltags(tags[sumleg], tags[ns[1]], nt[sumleg], nt[ns[1]]);
            if (p1 == 0 && nt[ns[1]] != 0) return false;
            if (su2_) {
                if (nt[ns[0]] * nt[ns[1]] == 0) {
                    *val = 1;
                } else {
                    const int twoJ = tags[sumleg][nt[sumleg] * BT - 1];
                    *val = p0 > p1 ? (twoJ == 0 ? 1 : -1) : 1;
                    if (sumleg == 1) *val *= twoJ == 0 ? -1 : 1;
                }
            } else {
                *val = nt[ns[0]] * nt[ns[1]] == 0 ? 1 : (p0 < p1 ? 1 : -1);
            }
            return true;
        }
        throw std::runtime_error("compare_tags: unexpected tag count");
    }
    bool correct_tags_4p(const int* const tags[3], const int nt[3], const int* order[4], int newpos[4]) const {
        // src/hamiltonian_qc.c:688-710
        int ncrean[2] = {0, 0};
        int cnt = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < nt[i]; ++j, ++cnt) {
                const int iscre = tags[i][j * BT];
                if (ncrean[iscre] == 2) return false;
                newpos[cnt] = iscre ? ncrean[iscre] : 2 + ncrean[iscre];
                order[newpos[cnt]] = &tags[i][j * BT];
                ++ncrean[iscre];
            }
        return true;
    }
    bool fuse_value(const int* const tags[3], const int nt[3], double* val) const {
        // src/hamiltonian_qc.c:497-576
        const int* order[4];
        int newpos[4];
        if (!correct_tags_4p(tags, nt, order, newpos)) return false;
        if (su2_) {
            *val = 1;
            if (nt[0] == 4 || nt[1] == 4 || nt[2] == 4) {
                *val = Bv(order, 0);
            } else if (nt[0] == 3 || nt[1] == 3 || nt[2] == 3) {
                *val = Bv(order, 0);
            } else {
                int twoJ = -1;
                int usetilde = 0;
                for (int l = 0; l < 3; ++l) {
                    if (nt[l] != 2) continue;
                    usetilde = tags[l][0] != tags[l][BT];
                    if (twoJ != -1 && twoJ != tags[l][2 * BT - 1]) return false;
                    twoJ = tags[l][2 * BT - 1];
                }
                if (nt[0] == 0 || nt[1] == 0 || nt[2] == 0) {
                    *val = usetilde ? Bt(order, twoJ) : Bv(order, twoJ);
                } else {
                    const int twoleg = (nt[1] == 2) + 2 * (nt[2] == 2);
                    const int tl[3][2] = {{1, 2}, {0, 2}, {0, 1}};
                    const int* nts[3] = {tags[twoleg], tags[tl[twoleg][0]], tags[tl[twoleg][1]]};
                    const int nnt[3] = {nt[twoleg], nt[tl[twoleg][0]], nt[tl[twoleg][1]]};
                    correct_tags_4p(nts, nnt, order, newpos);
                    if (usetilde) {
                        *val = Bt(order, twoJ);
                        if (tags[tl[twoleg][0]][0] != 1) *val *= twoJ == 2 ? -1 : 1;
                    } else {
                        *val = Bv(order, twoJ);
                    }
                    if (twoleg == 1) *val *= twoJ == 0 ? -1 : 1;
                }
            }
        } else {
            *val = 1;
            for (int i = 0; i < 4; ++i) {
                int j;
                for (j = i; j < 4; ++j)
                    if (newpos[j] == i) break;
                for (; j > i; --j) {
                    newpos[j] = newpos[j - 1];
                    *val *= -1;
                }
                newpos[i] = i;
            }
            *val *= get_V(order[0], order[1], order[2], order[3]) - get_V(order[0], order[1], order[3], order[2]);
        }
        return true;
    }

    // ---------------------------------------------------------- opTypes
    int symsec_tag(const int* tag, int nt) const {
        // src/hamiltonian_qc.c:1280-1302, 1378-1400
        const int n = su2_ ? 8 : 13;
        int hss[2] = {0, 0};
        int pgn = 0;
        if (su2_) {
            if (nt == 0) return trivial_hss();
            hss[1] = tag[nt * BT - 1];
            for (int i = 0; i < nt; ++i) {
                hss[0] += tag[i * BT] ? 1 : -1;
                pgn ^= orbirrep_[tag[i * BT + 1]];
            }
        } else {
            for (int i = 0; i < nt; ++i) {
                const int sg = tag[i * BT] ? 1 : -1;
                pgn ^= orbirrep_[tag[i * BT + 1]];
                hss[tag[i * BT + 2]] += sg;
            }
        }
        for (int i = 0; i < n; ++i)
            if (hss[0] == irr(i)[0] && hss[1] == irr(i)[1]) return i * npg_ + pgn;
        throw std::runtime_error("QC: no MPO sector for operator tag");
    }
    int symsec_of(const OpType& o, int nr, int typ, int k) const {
        // src/opType.c:553-563
        const int nt = kNrBaseTags[nr][typ];
        const int s = symsec_tag(o.tag(nr, typ, k) ? o.tag(nr, typ, k) : nullptr, nt < 0 ? 0 : nt);
        return typ == 1 ? hermitian_hss(s) : s;
    }
    std::vector<int> symsecs_of_ops(const OpType& o) const {
        std::vector<int> r;
        for (int i = 0; i < NR_OPS; ++i)
            for (int j = 0; j < NR_TYP; ++j)
                for (int k = 0; k < o.amount(i, j); ++k) r.push_back(symsec_of(o, i, j, k));
        return r;
    }

    bool need_double_ops(int bond, int is_left) const {
        const int l = netw_->nr_left_psites[bond], r = netw_->psites - l;
        return !(is_left == (l > r));
    }
    bool site_left_of_bond(int orb, int bond) const {
        const int* a = netw_->order_psites_of(bond, true);
        for (int i = 0; i < netw_->nr_left_psites[bond]; ++i)
            if (a[i] == orb) return true;
        return false;
    }
    bool need_compl_ops(int n, const int* ps, int bond, int is_left) const {
        // src/opType_qc.c:188-240
        if (n == 1 || n == 0) return true;
        const Network& N = *netw_;
        const int site = N.bonds[bond][is_left];
        const int l = N.nr_left_psites[bond], r = N.psites - l;
        const int prev = is_left ? l : r;
        if (is_left == (l > r)) return true;
        if (N.is_psite(site)) return false;
        int b[3], cnt = 0;
        N.bonds_of_site(site, b);
        for (int i = 0; i < 3; ++i)
            if (b[i] != bond) b[cnt++] = b[i];
        if (site_left_of_bond(ps[0], b[0]) != site_left_of_bond(ps[1], b[0])) return true;
        cnt = site_left_of_bond(ps[0], b[0]) ? 0 : 1;
        const int pb = cnt == 0 ? N.nr_left_psites[b[0]]
                                : (is_left ? N.psites - N.nr_left_psites[b[1]] : N.nr_left_psites[b[1]]);
        if (pb == prev) return b[cnt] > bond;
        return prev > pb;
    }

    // Unfiltered operator set of a bond (src/opType.c:423-432, 495-600).
    OpType init_optype(int bond, int is_left) const {
        const Network& N = *netw_;
        const int l = N.nr_left_psites[bond], r = N.psites - l;
        const int back = is_left ? l : r, forw = is_left ? r : l;
        const int* bl = N.order_psites_of(bond, is_left);
        const int* fl = N.order_psites_of(bond, !is_left);
        int count[NR_OPS][NR_TYP] = {{0}};
        OpType o;
        count[NR_OPS - 1][NR_TYP - 1] = back != 0;
        count[0][0] = 1;
        if (back != 0 && forw != 0) {
            for (int i = 0; i < NR_OPS; ++i)
                for (int j = 0; j < NR_TYP; ++j) {
                    const int nr = kNrBaseTags[i][j];
                    const bool compl_ = j == 1;
                    for (int kk = 0; kk < kTodo[i][j]; ++kk) {
                        const int* cr = kCreator[i][j][kk];
                        const int ns = compl_ ? forw : back;
                        const int* list = compl_ ? fl : bl;
                        auto emit = [&](const int* pos) {
                            // loop_dof (src/opType_qc.c:83-140)
                            const bool half = nr == 2 && cr[0] == cr[1] && pos[0] == pos[1];
                            if (!compl_ && nr == 2 && !need_double_ops(bond, is_left)) return;
                            if (compl_ && !need_compl_ops(nr, pos, bond, is_left)) return;
                            const int cas = (int)su2_ + (nr - 1 + half) * 2;
                            static const int maxpos[6] = {2, 1, 4, 2, 1, 1};
                            static const int q2[4][2] = {{0, 1}, {0, 0}, {1, 0}, {1, 1}};
                            for (int p = 0; p < maxpos[cas]; ++p) {
                                int dof[2] = {0, 0};
                                switch (cas) {
                                    case 0: dof[0] = p; break;
                                    case 1: dof[0] = 1; break;
                                    case 2:
                                    case 4: dof[0] = q2[p][0]; dof[1] = q2[p][1]; break;
                                    default: dof[0] = -1; dof[1] = 2 * p; break;
                                }
                                for (int t = 0; t < nr; ++t) {
                                    o.tags[i][j].push_back(cr[t]);
                                    o.tags[i][j].push_back(pos[t]);
                                    o.tags[i][j].push_back(dof[t]);
                                }
                                ++count[i][j];
                            }
                        };
                        if (ns == 0) continue;
                        if (nr == 1) {
                            for (int a = 0; a < ns; ++a) emit(&list[a]);
                        } else {
                            const bool hc = cr[0] == cr[1];
                            for (int a = 0; a < ns; ++a)
                                for (int b = hc ? a : 0; b < ns; ++b) {
                                    const int pos[2] = {list[a], list[b]};
                                    emit(pos);
                                }
                        }
                    }
                }
        }
        o.begin[0] = 0;
        for (int i = 0; i < NR_OPS; ++i)
            for (int j = 0; j < NR_TYP; ++j) o.begin[i * NR_TYP + j + 1] = o.begin[i * NR_TYP + j] + count[i][j];
        // sort each class by its tag tuple, last entry most significant
        for (int i = 0; i < NR_OPS; ++i)
            for (int j = 0; j < NR_TYP; ++j) {
                const int w = kNrBaseTags[i][j] * BT;
                const int n = o.amount(i, j);
                if (w <= 0 || n == 0) continue;
                std::vector<int> idx(n);
                for (int a = 0; a < n; ++a) idx[a] = a;
                const std::vector<int>& T = o.tags[i][j];
                std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
                    for (int k = w - 1; k >= 0; --k) {
                        const int x = T[(size_t)a * w + k], y = T[(size_t)b * w + k];
                        if (x != y) return x < y;
                    }
                    return false;
                });
                std::vector<int> nt((size_t)n * w);
                for (int a = 0; a < n; ++a)
                    std::copy(T.begin() + (size_t)idx[a] * w, T.begin() + (size_t)(idx[a] + 1) * w,
                              nt.begin() + (size_t)a * w);
                o.tags[i][j] = std::move(nt);
            }
        return o;
    }

    // get_opType: the filtered set once built, else the unfiltered one.
    OpType optype(int bond, int is_left, bool use_filtered) const {
        if (use_filtered) {
            auto it = filtered_.find({bond, is_left});
            if (it != filtered_.end()) return it->second;
        }
        return init_optype(bond, is_left);
    }
    OpType site_optype(int orb) const {
        // src/hamiltonian_qc.c:385-414 and change_site, src/opType.c:393-410
        OpType o;
        const int* nr = su2_ ? kSiteNrSU2 : kSiteNrU1;
        const int(*tg)[3] = su2_ ? kSiteTagsSU2 : kSiteTagsU1;
        int t = 0;
        o.begin[0] = 0;
        for (int i = 0; i < NR_OPS; ++i) {
            o.begin[i * 2 + 1] = o.begin[i * 2] + nr[i];
            o.begin[i * 2 + 2] = o.begin[i * 2 + 1];
            for (int j = 0; j < nr[i] * i; ++j, ++t) {
                o.tags[i][0].push_back(tg[t][0]);
                o.tags[i][0].push_back(orb);
                o.tags[i][0].push_back(tg[t][2]);
            }
        }
        return o;
    }
    static OpType unity_optype() {
        OpType o;
        for (int i = 1; i <= NR_OPS * NR_TYP; ++i) o.begin[i] = 1;
        return o;
    }
    static int id_unity(const OpType& o) { return o.amount(0, 0) ? 0 : -1; }
    static int id_H(const OpType& o) {
        return o.amount(NR_OPS - 1, 1) ? o.begin[(NR_OPS - 1) * NR_TYP + 1] : -1;
    }

    bool opType_exist(const OpType& o, const int* tag) const {
        // binary search in the sorted 2-'c' class (src/opType.c:140-157)
        const int K = o.amount(2, 1);
        const int w = 2 * BT;
        const std::vector<int>& T = o.tags[2][1];
        int lo = 0, hi = K - 1;
        while (lo <= hi) {
            const int mid = (lo + hi) / 2;
            int c = 0;
            for (int k = w - 1; k >= 0 && c == 0; --k) c = tag[k] - T[(size_t)mid * w + k];
            if (c == 0) return true;
            if (c < 0) hi = mid - 1;
            else lo = mid + 1;
        }
        return false;
    }

    bool interactval(const int ids[3], const OpType ops[3], char t, double* val) const {
        // src/opType_qc.c:163-186, 242-334
        int nr[3], typ[3], k[3], nt[3];
        const int* tags[3];
        for (int i = 0; i < 3; ++i) {
            ops[i].type_of(ids[i], &nr[i], &typ[i], &k[i]);
            nt[i] = kNrBaseTags[nr[i]][typ[i]];
            tags[i] = ops[i].tag(nr[i], typ[i], k[i]);
        }
        if (t == 't' || t == 'd') {
            if (typ[0] + typ[1] + typ[2] != 1) return false;
            const int cleg = (typ[1] == 1) + 2 * (typ[2] == 1);
            return compare_tags(tags, nt, cleg, val, t == 'd');
        }
        const int out = t - '1';
        const int st = typ[0] + typ[1] + typ[2];
        if (st == 0) return compare_tags(tags, nt, out, val, false);
        if (st == 1 && typ[out] == 1) {
            if (nt[out] == 0 || nt[out] == 1) {
                for (int l = 0; l < 3; ++l)
                    if (nt[l] == 2 && ops[l].amount(2, 1)) return false;
            } else if (nt[out] == 2) {
                for (int l = 0; l < 3; ++l)
                    if (nt[l] == 2 && out != l && opType_exist(ops[l], tags[out])) return false;
            }
            return fuse_value(tags, nt, val);
        }
        if (st == 2 && typ[out] == 1) {
            const int oc = (typ[0] == 1 && out != 0) * 0 + (typ[1] == 1 && out != 1) * 1 +
                           (typ[2] == 1 && out != 2) * 2;
            return compare_tags(tags, nt, oc, val, false);
        }
        return false;
    }

    void combine_all(const OpType ops[3], char c, const int order[3], InstructionSet* s) const {
        // src/instructions_qc.c:178-211, in sequential enumeration order
        for (int ci = 0; ci < 15; ++ci) {
            int opn[3] = {kCombine[ci][0], kCombine[ci][1], kCombine[ci][2]};
            if (c != 't' && c != 'd') opn[c - '1'] = 4 - opn[c - '1'];
            int off[3], amt[3];
            for (int j = 0; j < 3; ++j) ops[j].range(opn[j], &off[j], &amt[j]);
            for (int i2 = 0; i2 < amt[2]; ++i2)
                for (int i1 = 0; i1 < amt[1]; ++i1)
                    for (int i0 = 0; i0 < amt[0]; ++i0) {
                        const int cur[3] = {off[0] + i0, off[1] + i1, off[2] + i2};
                        double val;
                        if (!interactval(cur, ops, c, &val)) continue;
                        if (std::fabs(val) < 1e-20) continue;
                        s->instr.push_back(Instruction{{cur[order[0]], cur[order[1]], cur[order[2]]}, val});
                    }
        }
    }

    InstructionSet fetch_pupdate(const Network& n, int bond, int is_left, bool filtered) const {
        // src/instructions_qc.c:213-249
        InstructionSet s;
        const int site = n.bonds[bond][is_left];
        const int orb = n.sitetoorb[site];
        int b[3];
        n.bonds_of_site(site, b);
        OpType ops[3] = {optype(b[0], is_left, filtered), site_optype(orb), optype(b[2], is_left, filtered)};
        const int order[3] = {2 * !is_left, 1, 2 * is_left};
        s.step = 3;
        combine_all(ops, is_left ? '3' : '1', order, &s);
        if (b[0] == 0 && is_left) {
            const int ni[3] = {id_unity(ops[0]), id_unity(ops[1]), id_H(ops[2])};
            s.instr.push_back(Instruction{{ni[0], ni[1], ni[2]}, core_});
        }
        s.hss_of_new = symsecs_of_ops(optype(b[2 * is_left], is_left, filtered));
        return s;
    }
    InstructionSet fetch_bupdate(const Network& n, int bond, int is_left, bool filtered) const {
        // src/instructions_qc.c:251-282
        InstructionSet s;
        int b[3];
        n.bonds_of_site(n.bonds[bond][!is_left], b);
        const int uc = !is_left ? (b[1] == bond) : 2;
        const int bsite = n.bonds[bond][n.is_psite(n.bonds[bond][0])];
        static const int order[3][3] = {{1, 2, 0}, {0, 2, 1}, {0, 1, 2}};
        n.bonds_of_site(bsite, b);
        OpType ops[3] = {optype(b[0], uc != 0, filtered), optype(b[1], uc != 1, filtered),
                         optype(b[2], uc == 2, filtered)};
        s.step = 3;
        combine_all(ops, (char)('1' + uc), order[uc], &s);
        s.hss_of_new = symsecs_of_ops(optype(bond, is_left, filtered));
        return s;
    }

    void init_optype_array() {
        // Filtered operator sets of every bond, left then right
        // (src/opType.c:280-310, 412-447): keep the operators the
        // renormalisation actually produces.
        filtered_.clear();
        const Network& N = *netw_;
        for (int is_left = 1; is_left >= 0; --is_left)
            for (int i = 0; i < N.nr_bonds - 1; ++i) {
                const int bond = is_left ? i : N.nr_bonds - 1 - i;
                const int site = N.bonds[bond][!is_left];
                if (site == -1) {
                    filtered_[{bond, is_left}] = init_optype(bond, is_left);
                    continue;
                }
                InstructionSet s;
                if (N.is_psite(site)) {
                    int b[3];
                    N.bonds_of_site(site, b);
                    s = fetch_pupdate(N, b[2 * !is_left], is_left, true);
                } else {
                    s = fetch_bupdate(N, bond, is_left, true);
                }
                OpType full = optype(bond, is_left, true);
                const int ntot = full.begin[NR_OPS * NR_TYP];
                std::vector<char> keep(ntot, 0);
                for (const auto& in : s.instr)
                    if (in.instr[2] >= 0) keep[in.instr[2]] = 1;
                OpType f;
                int p = 0;
                f.begin[0] = 0;
                for (int a = 0; a < NR_OPS; ++a)
                    for (int t = 0; t < NR_TYP; ++t) {
                        const int cnt = full.amount(a, t);
                        const int w = std::max(0, kNrBaseTags[a][t]) * BT;
                        int kept = 0;
                        for (int k = 0; k < cnt; ++k, ++p) {
                            if (!keep[p]) continue;
                            ++kept;
                            if (w)
                                f.tags[a][t].insert(f.tags[a][t].end(), full.tags[a][t].begin() + (size_t)k * w,
                                                    full.tags[a][t].begin() + (size_t)(k + 1) * w);
                        }
                        f.begin[a * NR_TYP + t + 1] = f.begin[a * NR_TYP + t] + kept;
                    }
                filtered_[{bond, is_left}] = std::move(f);
            }
    }

    // ----------------------------------------------------- site elements
    static double u1_el(int op, int bra, int ket) {
        // src/hamiltonian_qc.c:1211-1278
        switch (op) {
            case 0: return bra == ket ? 1.0 : 0.0;
            case 1: return (bra == 1 && ket == 0) ? 1.0 : (bra == 3 && ket == 2) ? -1.0 : 0.0;
            case 2: return (bra == 2 && ket == 0) ? 1.0 : (bra == 3 && ket == 1) ? 1.0 : 0.0;
            case 3: return (bra == 0 && ket == 1) ? -1.0 : (bra == 2 && ket == 3) ? 1.0 : 0.0;
            case 4: return (bra == 0 && ket == 2) ? -1.0 : (bra == 1 && ket == 3) ? -1.0 : 0.0;
            case 5: return (bra == 3 && ket == 0) ? 1.0 : 0.0;
            case 6: return (bra == 0 && ket == 3) ? -1.0 : 0.0;
            case 7: return ((bra == 1 && ket == 1) || (bra == 3 && ket == 3)) ? 1.0 : 0.0;
            case 8: return (bra == 2 && ket == 1) ? 1.0 : 0.0;
            case 9: return (bra == 1 && ket == 2) ? 1.0 : 0.0;
            case 10: return ((bra == 2 && ket == 2) || (bra == 3 && ket == 3)) ? 1.0 : 0.0;
            case 11: return (bra == 3 && ket == 1) ? -1.0 : 0.0;
            case 12: return (bra == 3 && ket == 2) ? -1.0 : 0.0;
            case 13: return (bra == 1 && ket == 3) ? -1.0 : 0.0;
            case 14: return (bra == 2 && ket == 3) ? -1.0 : 0.0;
            case 15: return (bra == 3 && ket == 3) ? -1.0 : 0.0;
            default: throw std::runtime_error("QC: bad site operator");
        }
    }
    static double su2_el(int op, int bra, int ket) {
        // src/hamiltonian_qc.c:1316-1376
        const double s2 = std::sqrt(2.0);
        switch (op) {
            case 0: return bra == ket ? (bra == 1 ? -s2 : 1.0) : 0.0;
            case 1: return ((bra == 1 && ket == 0) || (bra == 2 && ket == 1)) ? s2 : 0.0;
            case 2: return (bra == 0 && ket == 1) ? -s2 : (bra == 1 && ket == 2) ? s2 : 0.0;
            case 3: return (bra == 2 && ket == 0) ? 1 / s2 : 0.0;
            case 4: return (bra == 0 && ket == 2) ? -1 / s2 : 0.0;
            case 5: return (bra == 1 && ket == 1) ? -1.0 : (bra == 2 && ket == 2) ? s2 : 0.0;
            case 6: return (bra == 1 && ket == 1) ? -std::sqrt(3.0) : 0.0;
            case 7: return (bra == 2 && ket == 1) ? -1.0 : 0.0;
            case 8: return (bra == 1 && ket == 2) ? -1.0 : 0.0;
            case 9: return (bra == 2 && ket == 2) ? -0.5 : 0.0;
            default: throw std::runtime_error("QC: bad site operator");
        }
    }
};

}  // namespace

std::unique_ptr<Model> make_qc_model(const std::string& text, bool su2) {
    return std::unique_ptr<Model>(new QC(text, su2));
}

std::unique_ptr<Model> make_model_from_data(const HamData& d) {
    if (d.type == HAM_QC) return std::unique_ptr<Model>(new QC(d));
    if (d.type == HAM_NN_HUBBARD) return make_hubbard_model(d.t, d.U, d.su2 != 0);
    throw std::runtime_error("checkpoint: Hamiltonian type " + std::to_string(d.type) +
                             " is not supported");
}

}  // namespace t3
