// Symmetry algebra for block-sparse tree tensor networks: irrep products,
// fermionic Z2 signs and SU(2) recoupling prefactors, Wigner 6j/9j symbols.
//
// Follows the reference's symmetry layer:
//   include/symmetries.h:34 (group enum), src/symmetries.c:158-332 (dispatch),
//   src/symmetry_z2.c:54-228, src/symmetry_su2.c:26-120, src/symmetry_u1.c:35-49,
//   src/symmetry_pg.c:63-69, src/Wigner.c:230-330 (Racah formula).
#pragma once
#include <array>
#include <cstdint>
#include <string>

namespace t3 {

constexpr int MAX_SYMMETRIES = 5;
using QN = int64_t;
using Irrep = std::array<int, MAX_SYMMETRIES>;

// Same ordering as include/symmetries.h:34 with POINT_GROUP_SYMMETRY expanded
// (include/symmetry_pg.h:65).
enum SymGroup : int { Z2 = 0, U1, SU2, C1, Ci, C2, Cs, D2, C2v, C2h, D2h, SENIORITY };

struct SymInfo {
    int nrSyms = 0;
    int sgs[MAX_SYMMETRIES] = {0};
    int target_state[MAX_SYMMETRIES] = {0};
};

const char* symstring(int sg);
int symgroup_from_string(const std::string& s);  // -1 if unknown
int pg_nr_irreps(int sg);
int pg_irrep_from_string(int sg, const std::string& s);  // -1 if unknown
// Molpro (FCIDUMP) irrep label 1..n -> psi4 ordering index (src/symmetry_pg.c:41-89)
int fcidump_to_psi4(int fcidump_irrep0, int sg);

// Tensor product of two irreps: resulting irreps are min, min+step, ...
// (nr of them).  src/symmetries.c:42-66
void tensprod_irrep(int* min_irrep, int* nr_irreps, int* step, int irrep1,
                    int irrep2, int sign, int sg);

// Upper bound (exclusive) for the irrep label range of products.
// src/symmetries.c:27-40
int max_irrep_bound(int max1, int max2, int sg);

// Wigner symbols with 2j arguments (doubled spins).  Racah's closed form.
double wigner6j(int a, int b, int c, int d, int e, int f);
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);

// Prefactors.  Arguments mirror src/symmetries.c:158-332; the irreps are given
// as pointers to full irrep arrays and the symmetry index is selected inside.
double prefactor_pAppend(const int* const (*irr)[3], int is_left, const SymInfo& s);
double prefactor_adjoint(const int* const* irr, char c, const SymInfo& s);
double prefactor_pUpdate(const int* const (*irr)[3], int is_left, const SymInfo& s);
double prefactor_mirror_coupling(const int* const* irr, const SymInfo& s);
double prefactor_bUpdate(const int* const (*irr)[3], int updateCase, const SymInfo& s);
double prefactor_add_P_operator(const int* const (*irr)[3], int isleft, const SymInfo& s);
double prefactor_combine_MPOs(const int* const (*irr)[3], const int* const* irrMPO,
                              const SymInfo& s, int isdmrg, int extradinge);

bool need_multiplicity(const SymInfo& s);
int multiplicity(const SymInfo& s, const int* irrep);

}  // namespace t3
