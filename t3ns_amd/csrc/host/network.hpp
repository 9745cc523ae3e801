// Tree tensor network topology: sites, bonds, sweep order and the
// optimisation-step scheduler.  Follows src/network.c and include/network.h.
// The reference keeps this as the process global `netw` (src/network.c:27) and
// the scheduler state as a function static (src/network.c:734); here both are
// members of a Network object so several solves can coexist in one process.
#pragma once
#include <array>
#include <string>
#include <vector>

namespace t3 {

constexpr int STEPSPECS_MBONDS = 3;  // include/network.h:50
constexpr int STEPSPECS_MSITES = 4;  // include/network.h:52

struct StepSpecs {  // include/network.h:55-84
    int nr_bonds_opt = 0;
    int bonds_opt[STEPSPECS_MBONDS] = {-1, -1, -1};
    int nr_sites_opt = 0;
    int sites_opt[STEPSPECS_MSITES] = {-1, -1, -1, -1};
    int common_next[STEPSPECS_MSITES] = {0, 0, 0, 0};
    int nCenter = -1;
};

class Network {
public:
    int nr_bonds = 0;
    int psites = 0;
    int sites = 0;
    std::vector<std::array<int, 2>> bonds;
    std::vector<int> sitetoorb;
    std::vector<int> nr_left_psites;
    std::vector<std::vector<int>> order_psites;
    std::vector<int> sweep;

    // Parses the .netw text format (src/network.c:207-354).
    static Network from_text(const std::string& text);
    static Network from_file(const std::string& path);
    // Builds a network from an explicit bond list and site->orbital map.
    static Network from_bonds(int nsites, const std::vector<int>& sitetoorb,
                              const std::vector<std::array<int, 2>>& bonds);
    // The same with a given sweep order (empty: the default simple sweep).
    static Network from_bonds(int nsites, const std::vector<int>& sitetoorb,
                              const std::vector<std::array<int, 2>>& bonds,
                              const std::vector<int>& sweep);

    bool is_psite(int site) const { return sitetoorb[site] >= 0; }
    bool is_pbond(int bond) const { return bond >= 2 * nr_bonds; }
    // [in1, in2/phys, out] of a site (src/network.c:414-445)
    void bonds_of_site(int site, int* out) const;
    int common_bond(int site1, int site2) const;
    bool is_dmrg_bond(int bond) const;
    int outgoing_bond() const;
    const int* order_psites_of(int bond, bool is_left) const;

    // Optimisation-step scheduler (src/network.c:529-727).
    bool next_opt_step(int maxsites, StepSpecs* specs);
    void reset_sweep_state() { curr_state_ = 0; }

private:
    int curr_state_ = 0;
    void finalize();
    void make_simplesweep(bool inclborder);
    void recursive_stepping(bool inclborder, int* id);
    void move_forward_state(const StepSpecs* specs, int* state) const;
    bool get_sites_to_opt(int maxsites, StepSpecs* specs, int* state) const;
    void get_bonds_involved(StepSpecs* specs) const;
    void get_common_with_next(int maxsites, StepSpecs* specs, int next_state) const;
    void set_nCenter(StepSpecs* specs) const;
};

// Open chain of L physical sites (the DMRG network).
Network make_chain_network(int L);

}  // namespace t3
