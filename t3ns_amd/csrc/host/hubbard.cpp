// Nearest-neighbour Hubbard model: site operators, MPO sectors and the
// instruction lists for renormalisation and Heff merging.
// Follows src/hamiltonian_nn_hubbard.c:36-441 and
// src/instructions_nn_hubbard.c:28-284.
#include <cmath>
#include <stdexcept>

#include "model.hpp"

namespace t3 {
namespace {

// MPO-bond irreps: (dN_up, dN_down) for U1xU1, (dN, 2S) for SU2
// (src/hamiltonian_nn_hubbard.c:36-38).
const int kMpoU1[5][2] = {{-1, 0}, {0, -1}, {0, 0}, {1, 0}, {0, 1}};
const int kMpoSU2[3][2] = {{-1, 1}, {0, 0}, {1, 1}};

class Hubbard final : public Model {
public:
    HamData ham_data() const override {
        HamData d;
        d.type = HAM_NN_HUBBARD;
        d.su2 = su2_;
        d.t = t_;
        d.U = U_;
        return d;
    }
    Hubbard(double t, double U, bool su2, bool quirk) : t_(t), U_(U), su2_(su2), quirk_(quirk) {
        const int n = su2_ ? 3 : 5;
        mpo_.bond = -1;
        mpo_.nrSecs = n;
        mpo_.irreps.assign(n, Irrep{});
        mpo_.fcidims.assign(n, 1.0);
        mpo_.dims.assign(n, 1);
        mpo_.totaldims = n;
        for (int i = 0; i < n; ++i) {
            const int a = irr(i)[0], b = irr(i)[1];
            mpo_.irreps[i][0] = su2_ ? std::abs(a) % 2 : std::abs(a + b) % 2;
            mpo_.irreps[i][1] = a;
            mpo_.irreps[i][2] = b;
        }
    }
    std::string name() const override { return "NN_HUBBARD"; }

    Symsecs phys_symsecs(int orbital) const override {
        // Local states |0>, |up>, |down>, |updown> (src/hamiltonian_nn_hubbard.c:89-113)
        static const int u1[4][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
        static const int s2[3][3] = {{0, 0, 0}, {1, 1, 1}, {0, 2, 0}};
        Symsecs r;
        r.nrSecs = su2_ ? 3 : 4;
        r.totaldims = r.nrSecs;
        r.irreps.assign(r.nrSecs, Irrep{});
        r.dims.assign(r.nrSecs, 1);
        r.fcidims.assign(r.nrSecs, 1.0);
        for (int i = 0; i < r.nrSecs; ++i)
            for (int j = 0; j < 3; ++j) r.irreps[i][j] = su2_ ? s2[i][j] : u1[i][j];
        (void)orbital;
        return r;
    }
    const Symsecs& mpo_symsecs() const override { return mpo_; }
    int trivial_hss() const override { return su2_ ? 1 : 2; }
    int hermitian_hss(int h) const override {
        const int a = -irr(h)[0], b = irr(h)[1] * (su2_ ? 1 : -1);
        for (int i = 0; i < mpo_.nrSecs; ++i)
            if (irr(i)[0] == a && irr(i)[1] == b) return i;
        return -1;
    }
    bool mpo_couples_to_singlet(const int m[3]) const override { return valid_tprod(m[0], m[1], m[2]); }

    int symsec_siteop(int siteop, int) const override {
        int a, b;
        if (su2_) {
            switch (siteop) {
                case 1: a = 1; b = 1; break;
                case 2: a = -1; b = 1; break;
                case 0:
                case 8: a = 0; b = 0; break;
                default: throw std::runtime_error("hubbard: bad siteop");
            }
        } else {
            switch (siteop) {
                case 20: a = -1; b = 0; break;
                case 21: a = 0; b = -1; break;
                case 0:
                case 8: a = 0; b = 0; break;
                case 11: a = 0; b = 1; break;
                case 10: a = 1; b = 0; break;
                default: throw std::runtime_error("hubbard: bad siteop");
            }
        }
        for (int i = 0; i < mpo_.nrSecs; ++i)
            if (irr(i)[0] == a && irr(i)[1] == b) return i;
        return -1;
    }

    double el_siteop(int op, int bra, int ket) const override {
        // Matrix elements include the |ket(i)||MPO(i)| sign convention of the
        // reference (src/hamiltonian_nn_hubbard.c:395-441, 487-521).
        if (!su2_) {
            switch (op) {
                case 0: return bra == ket ? 1.0 : 0.0;
                case 10:
                    if (bra == 1 && ket == 0) return 1.0;
                    if (bra == 3 && ket == 2) return -1.0;
                    return 0.0;
                case 11:
                    if (bra == 2 && ket == 0) return 1.0;
                    if (bra == 3 && ket == 1) return 1.0;
                    return 0.0;
                case 20:
                    if (bra == 0 && ket == 1) return -1.0;
                    if (bra == 2 && ket == 3) return 1.0;
                    return 0.0;
                case 21:
                    if (bra == 0 && ket == 2) return -1.0;
                    if (bra == 1 && ket == 3) return -1.0;
                    return 0.0;
                case 8: return (bra == 3 && ket == 3) ? -1.0 : 0.0;
                default: throw std::runtime_error("hubbard: bad siteop");
            }
        }
        const double s2 = std::sqrt(2.0);
        switch (op) {
            case 0:
                if (bra == ket) return bra == 1 ? -s2 : 1.0;
                return 0.0;
            case 1:
                if (bra == 1 && ket == 0) return s2;
                if (bra == 2 && ket == 1) return s2;
                return 0.0;
            case 2:
                if (bra == 0 && ket == 1) return -s2;
                if (bra == 1 && ket == 2) return s2;
                return 0.0;
            case 8: return (bra == 2 && ket == 2) ? 1.0 : 0.0;
            default: throw std::runtime_error("hubbard: bad siteop");
        }
    }

    std::vector<std::array<int, 2>> tprods_ham(int res, int) const override {
        std::vector<std::array<int, 2>> out;
        for (int i = 0; i < mpo_.nrSecs; ++i)
            for (int j = 0; j < mpo_.nrSecs; ++j)
                if (valid_tprod(i, j, res)) out.push_back({i, j});
        return out;
    }

    bool consistent_state(const int* ts, const SymInfo& s) const override {
        // src/hamiltonian_nn_hubbard.c:289-317
        int parity = 0, particles = 0, su2v = -1;
        for (int i = 0; i < s.nrSyms; ++i) {
            switch (s.sgs[i]) {
                case Z2: parity = ts[i]; break;
                case U1:
                    particles += ts[i];
                    if (ts[i] > s.target_state[i]) return false;
                    break;
                case SU2: su2v = ts[i]; break;
                default: throw std::runtime_error("invalid symmetry for NN_HUBBARD");
            }
        }
        if (parity != particles % 2) return false;
        if (su2v != -1 && particles % 2 != su2v % 2) return false;
        return true;
    }

    InstructionSet raw_pUpdate(const Network& n, int bond, int is_left) override {
        const bool border = n.bonds[bond][!is_left] == -1;
        InstructionSet s;
        s.step = 3;
        if (su2_) {
            const double t = t_ * std::sqrt(2.0);
            add(s, 0, 0, 0, 1);
            add(s, 0, 1, 1, 1);
            add(s, 0, 2, 2, 1);
            add(s, 0, 8, 3, U_);
            if (!border) {
                add(s, 1, 2, 3, -t);
                add(s, 2, 1, 3, -t);
                add(s, 3, 0, 3, 1);
            }
            s.hss_of_new = {symsec_siteop(0, 0), symsec_siteop(1, 0), symsec_siteop(2, 0),
                            symsec_siteop(8, 0)};
        } else {
            const double sign = is_left ? 1 : -1;
            add(s, 0, 0, 0, 1);
            add(s, 0, 10, 1, 1);
            add(s, 0, 11, 2, 1);
            add(s, 0, 20, 3, 1);
            add(s, 0, 21, 4, 1);
            add(s, 0, 8, 5, -U_);
            if (!border) {
                add(s, 1, 20, 5, -sign * t_);
                add(s, 2, 21, 5, -sign * t_);
                add(s, 3, 10, 5, sign * t_);
                add(s, 4, 11, 5, sign * t_);
                add(s, 5, 0, 5, 1);
            }
            s.hss_of_new = hss_u1();
        }
        return s;
    }

    InstructionSet raw_bUpdate(const Network& n, int bond, int is_left) override {
        int bz[3];
        n.bonds_of_site(n.bonds[bond][!is_left], bz);
        const int uc = !is_left ? (bz[1] == bond) : 2;
        InstructionSet s;
        s.step = 3;
        if (su2_) {
            const double t = t_ * std::sqrt(2.0);
            add(s, 0, 0, 0, 1);
            add(s, 0, 1, 1, uc == 0 ? -1 : 1);
            add(s, 0, 2, 2, uc == 0 ? -1 : 1);
            add(s, 0, 3, 3, 1);
            add(s, 1, 0, 1, uc == 2 ? -1 : 1);
            add(s, 1, 2, 3, uc == 1 ? t : -t);
            add(s, 2, 0, 2, uc == 2 ? -1 : 1);
            add(s, 2, 1, 3, uc == 1 ? t : -t);
            add(s, 3, 0, 3, 1);
            s.hss_of_new = {symsec_siteop(0, 0), symsec_siteop(1, 0), symsec_siteop(2, 0),
                            symsec_siteop(8, 0)};
        } else {
            const double sg = uc == 1 ? -1 : 1;
            for (int i = 0; i < 6; ++i) add(s, 0, i, i, 1);
            add(s, 1, 0, 1, sg);
            add(s, 2, 0, 2, sg);
            add(s, 3, 0, 3, sg);
            add(s, 4, 0, 4, sg);
            add(s, 5, 0, 5, 1);
            add(s, 1, 3, 5, -t_);
            add(s, 2, 4, 5, -t_);
            add(s, 3, 1, 5, t_);
            add(s, 4, 2, 5, t_);
            s.hss_of_new = hss_u1();
        }
        return s;
    }

    InstructionSet raw_merge(const Network&, int bond, int isdmrg) override {
        const bool dm = quirk_ ? bond != 0 : isdmrg != 0;
        InstructionSet s;
        if (su2_) {
            const double t = t_ * std::sqrt(2.0);
            if (dm) {
                s.step = 2;
                add(s, 0, 3, 0, 1);
                add(s, 1, 2, 0, -t);
                add(s, 2, 1, 0, -t);
                add(s, 3, 0, 0, 1);
            } else {
                s.step = 3;
                add(s, 0, 0, 3, 1);
                add(s, 0, 1, 2, -t);
                add(s, 0, 2, 1, -t);
                add(s, 0, 3, 0, 1);
                add(s, 1, 0, 2, t);
                add(s, 2, 0, 1, t);
                add(s, 1, 2, 0, -t);
                add(s, 2, 1, 0, -t);
                add(s, 3, 0, 0, 1);
            }
        } else {
            if (dm) {
                s.step = 2;
                add(s, 0, 5, 0, 1);
                add(s, 1, 3, 0, -t_);
                add(s, 2, 4, 0, -t_);
                add(s, 3, 1, 0, t_);
                add(s, 4, 2, 0, t_);
                add(s, 5, 0, 0, 1);
            } else {
                s.step = 3;
                add(s, 0, 0, 5, 1);
                add(s, 0, 1, 3, -t_);
                add(s, 0, 2, 4, -t_);
                add(s, 0, 3, 1, t_);
                add(s, 0, 4, 2, t_);
                add(s, 0, 5, 0, 1);
                add(s, 1, 0, 3, -t_);
                add(s, 2, 0, 4, -t_);
                add(s, 3, 0, 1, t_);
                add(s, 4, 0, 2, t_);
                add(s, 5, 0, 0, 1);
                add(s, 1, 3, 0, -t_);
                add(s, 2, 4, 0, -t_);
                add(s, 3, 1, 0, t_);
                add(s, 4, 2, 0, t_);
            }
        }
        return s;
    }

private:
    double t_, U_;
    bool su2_, quirk_;
    Symsecs mpo_;

    const int* irr(int i) const { return su2_ ? kMpoSU2[i] : kMpoU1[i]; }
    bool valid_tprod(int i, int j, int r) const {
        const int* a = irr(i);
        const int* b = irr(j);
        const int* c = irr(r);
        if (su2_)
            return a[0] + b[0] == c[0] && (a[1] + b[1] + c[1]) % 2 == 0 && c[1] <= a[1] + b[1] &&
                   c[1] >= std::abs(a[1] - b[1]);
        return a[0] + b[0] == c[0] && a[1] + b[1] == c[1];
    }
    std::vector<int> hss_u1() const {
        return {symsec_siteop(0, 0), symsec_siteop(10, 0), symsec_siteop(11, 0),
                symsec_siteop(20, 0), symsec_siteop(21, 0), symsec_siteop(8, 0)};
    }
    static void add(InstructionSet& s, int a, int b, int c, double p) {
        s.instr.push_back(Instruction{{a, b, c}, p});
    }
};

}  // namespace

std::unique_ptr<Model> make_hubbard_model(double t, double U, bool su2, bool quirk) {
    return std::unique_ptr<Model>(new Hubbard(t, U, su2, quirk));
}

}  // namespace t3
