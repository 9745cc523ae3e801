// Symmetry sectors, instruction sets and the Hamiltonian (MPO) model
// interface.  The reference dispatches on a process-global `ham` enum
// (src/hamiltonian.c:31); here a Model object carries the same operations
// (include/hamiltonian.h:30-120, include/instructions.h:39-86).
#pragma once
#include <array>
#include <memory>
#include <string>
#include <vector>

#include "network.hpp"
#include "sym.hpp"

namespace t3 {

// Symmetry sectors of one bond (include/symsecs.h:32-54).
struct Symsecs {
    int bond = 0;
    int nrSecs = 0;
    std::vector<Irrep> irreps;
    std::vector<double> fcidims;
    std::vector<int> dims;
    int totaldims = 0;
};

struct Instruction {  // include/instructions.h:24-33
    int instr[3];
    double pref;
};

struct InstructionSet {  // include/instructions.h:36-66
    int step = 0;
    std::vector<Instruction> instr;
    std::vector<int> hss_of_new;
    // Merge sets only: MPO combination per group and group starts.
    std::vector<int> MPOc;
    std::vector<int> MPOc_beg;
    int nr_instr() const { return (int)instr.size(); }
};

// Hamiltonian data as the reference checkpoints it (src/hamiltonian.c:
// 327-345, src/hamiltonian_qc.c:640-654, src/hamiltonian_nn_hubbard.c:
// 261-270).
struct HamData {
    int type = 0;                // enum hamtypes (include/hamiltonian.h:30): 1 QC, 2 NN_HUBBARD
    int su2 = 0;
    int norb = 0;                // QC
    std::vector<int> orbirrep;   // QC: psi4 irrep of every orbital
    std::vector<double> Vijkl;   // QC: norb^4, the one-body part folded in
    double core_energy = 0;      // QC
    int has_seniority = 0;       // QC
    double t = 0, U = 0;         // NN_HUBBARD
};
enum HamType { HAM_QC = 1, HAM_NN_HUBBARD = 2 };

class Model {
public:
    virtual ~Model() = default;
    virtual std::string name() const = 0;
    // What a checkpoint stores of the Hamiltonian.
    virtual HamData ham_data() const = 0;
    virtual Symsecs phys_symsecs(int orbital) const = 0;
    virtual const Symsecs& mpo_symsecs() const = 0;
    int nr_hss() const { return mpo_symsecs().nrSecs; }
    virtual int trivial_hss() const = 0;
    virtual int hermitian_hss(int hss) const = 0;
    virtual bool mpo_couples_to_singlet(const int mpo[3]) const = 0;
    // Site operator `siteop` at network site `site`: its MPO sector and its
    // matrix elements.
    virtual int symsec_siteop(int siteop, int site) const = 0;
    virtual double el_siteop(int siteop, int braid, int ketid) const = 0;
    // All (hss_i, hss_j) whose product can give `resulting` at network site
    // `site` (src/hamiltonian.c:227-251).
    virtual std::vector<std::array<int, 2>> tprods_ham(int resulting, int site) const = 0;
    virtual bool consistent_state(const int* ts, const SymInfo& s) const = 0;
    // Called once the network and symmetries are known.
    virtual void attach(const Network& n, const SymInfo& s) { (void)n; (void)s; }
    // Raw (unsorted) instruction lists as the model generates them.
    virtual InstructionSet raw_pUpdate(const Network& n, int bond, int is_left) = 0;
    virtual InstructionSet raw_bUpdate(const Network& n, int bond, int is_left) = 0;
    virtual InstructionSet raw_merge(const Network& n, int bond, int isdmrg) = 0;
};

// NN_HUBBARD (t, U) model on any network; U1xU1 = (N_up, N_down) or
// U1xSU2 = (N, 2S) (src/hamiltonian_nn_hubbard.c, src/instructions_nn_hubbard.c).
// With reference_merge_quirk the merge set is chosen as the reference does
// (by `bond != 0`, src/instructions.c:204), otherwise by isdmrg.
std::unique_ptr<Model> make_hubbard_model(double t, double U, bool su2,
                                          bool reference_merge_quirk = false);

// Quantum chemistry model from FCIDUMP text (src/hamiltonian_qc.c,
// src/opType.c, src/opType_qc.c, src/instructions_qc.c).  su2 selects U1xSU2,
// otherwise U1xU1; a point group is taken from the symmetry list at attach().
std::unique_ptr<Model> make_qc_model(const std::string& fcidump_text, bool su2);

// The model a checkpoint describes (QC from its stored integrals, which
// already hold the folded one-body part; NN_HUBBARD from t, U, su2).
std::unique_ptr<Model> make_model_from_data(const HamData& d);

}  // namespace t3
