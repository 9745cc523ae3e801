// Solve context: network, model, symmetry sectors of every bond (the
// reference's bookkeeper) and cached instruction sets.  The reference keeps
// these as process globals `bookie`, `netw`, `ham` and static instruction
// caches (src/bookkeeper.c:30, src/network.c:27, src/hamiltonian.c:31,
// src/instructions.c:33-37); a Context owns them so solves are re-entrant.
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "model.hpp"
#include "network.hpp"
#include "sym.hpp"

namespace t3 {

struct GSec {
    int d;    // dimension of the block
    int id3;  // sector index on the third bond
};

// Valid sector triples of a three-legged coupling (src/tensorproducts.c:114-142).
struct GoodSectors {
    const Symsecs* ss[3] = {nullptr, nullptr, nullptr};
    std::vector<char> rowvalid;                    // ss[0].dims[i] != 0
    std::vector<std::vector<std::vector<GSec>>> sec;  // [i][j] -> list
    int total = 0;
};

struct GsEntry {
    int cid[3];
    QN cqn;
    int cdim;
};

class Context {
public:
    SymInfo sym;
    Network netw;
    std::unique_ptr<Model> ham;
    std::vector<Symsecs> v_symsecs;
    std::vector<Symsecs> p_symsecs;
    // Whether the host layer's problem builders fill operators and tensors
    // with seeded values (off: structure only, elements zero, for callers
    // that supply their own data).
    bool fill_seeded = true;

    // Symmetry sectors of a bond index in the reference numbering:
    // [0, 2*nr_bonds) virtual (ket, bra), >= 2*nr_bonds physical, -1 MPO
    // (src/bookkeeper.c:792-815).
    const Symsecs& symsecs(int bond) const;
    Symsecs& symsecs_mut(int bond);
    int get_braT3NSbond(int bond) const;
    int get_ketT3NSbond(int bond) const;

    int search_symsec(const int* irrep, const Symsecs& ss) const;
    bool irreps_equal(const int* a, const int* b) const;

    // Builds v_symsecs/p_symsecs (preparebookkeeper with no previous state,
    // src/bookkeeper.c:725-776).
    void prepare_bookkeeper(int max_dim, bool interm_scale, int minocc);

    Symsecs tensprod_symsecs(const Symsecs& s1, const Symsecs& s2, int sign, char o) const;
    void kick_empty_symsecs(Symsecs& s, char o) const;
    GoodSectors find_good_sectors(const Symsecs* const ss[3], int sign) const;
    // Entries in the iteration order of iterate_gs (src/tensorproducts.c:144-249).
    std::vector<GsEntry> iter_gs(const GoodSectors& gs, int idnr, int tid) const;
    int iter_gs_length(const GoodSectors& gs, int idnr, int tid) const;

    // Sorted instruction sets, cached per (bond, is_left / isdmrg).
    const InstructionSet& fetch_pUpdate(int bond, int is_left);
    const InstructionSet& fetch_bUpdate(int bond, int is_left);
    const InstructionSet& fetch_merge(int bond, int isdmrg, const int* const* hss_ops);
    // Unsorted merge instructions of a bond with their flat (operator
    // triple, prefactor) arrays, cached per (bond, isdmrg): they depend on
    // the network and the Hamiltonian only.
    struct RawMerge {
        InstructionSet set;
        std::vector<int> instr;
        std::vector<double> pref;
    };
    const RawMerge& fetch_raw_merge(int bond, int isdmrg);

    bool bond_is_internal_dims1(int bond) const;
    int full_dimension(const Symsecs& s) const;

private:
    std::map<std::pair<int, int>, InstructionSet> pupd_, bupd_, merge_;
    std::map<std::pair<int, int>, RawMerge> rawmerge_;
    void make_symsec(Symsecs* symsec, int bond, int is_left, char o);
    bool select_lowest(Symsecs& s1, const Symsecs& s2) const;
    void kick_impossibles(Symsecs& s) const;
    Symsecs init_vacuum(char o) const;
    Symsecs init_target(char o) const;
    void build_all_sectors(Symsecs& res, const Symsecs& s1, const Symsecs& s2) const;
};

// Builds a context from a network, a model and the symmetry/target-state
// description; Z2 is prepended automatically as in include_Z2
// (src/bookkeeper.c:495-513).  Symmetries are sorted by group.
std::unique_ptr<Context> make_context(Network netw, std::unique_ptr<Model> ham,
                                      const std::vector<int>& sgs,
                                      const std::vector<int>& target_state);

// Sorting of an instruction list by (instr[0], instr[1]), stable
// (src/instructions.c:51-55 with the comparator of src/sort.c:92-99).
void sort_instructions(InstructionSet& s);
// Groups a merge set by MPO combination (src/instructions.c:57-88).
void sort_instructions_merge(InstructionSet& s, const int* const* hss_ops, int nrhss);
// Iterates unique (old1, old2, hss_new) instructions (src/instructions.c:224-251).
bool next_unique_instr(int* curr, const InstructionSet& s);

}  // namespace t3
