/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C (C11 + OpenMP) restatement of the reference's algorithms on the
 * Heff-matvec path, used only by tests/, the smoke check and the CPU-baseline
 * leg of bench.py as the checker.  Each function cites the reference
 * file:line it follows.  It takes the same boundary descriptors as
 * include/t3ns_amd.h and shares no code with the product library.
 *
 * Pinning (see tests/): symmetry prefactors and Wigner symbols against
 * oracle/_ref (the reference's own symmetry sources compiled unchanged) and
 * sympy; GEMM against numpy; the matvec against closed forms, Hermiticity,
 * diagonal = unit-vector matvec, and exact diagonalisation.
 */
#ifndef T3NS_ORACLE_H
#define T3NS_ORACLE_H

#include <stdint.h>

#include "../include/t3ns_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Symmetry helpers (src/symmetries.c:158-332, src/Wigner.c:245-330). */
double oracle_wigner6j(int a, int b, int c, int d, int e, int f);
double oracle_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);
/* irr: [2][3] irreps (NEW/OLD x bond) as flat int[2*3*5]; see symmetries.c:280 */
double oracle_prefactor_add_P_operator(const int* irr, int isleft, const int* sgs, int nrsy);
/* irr: [2][3][5], mpo: [3][5]; src/symmetries.c:301-332 */
double oracle_prefactor_combine_MPOs(const int* irr, const int* mpo, const int* sgs, int nrsy,
                                     int isdmrg, int extradinge);
/* irr: [3][3][5]; src/symmetries.c:158-183 */
double oracle_prefactor_pAppend(const int* irr, int is_left, const int* sgs, int nrsy);
/* irr: [3][5]; src/symmetries.c:185-201 */
double oracle_prefactor_adjoint(const int* irr, char c, const int* sgs, int nrsy);
/* irr: [3][3][5]; src/symmetries.c:213-235 */
double oracle_prefactor_pUpdate(const int* irr, int is_left, const int* sgs, int nrsy);
/* irr: [3][5]; src/symmetries.c:237-251 */
double oracle_prefactor_mirror_coupling(const int* irr, const int* sgs, int nrsy);
/* irr: [3][3][5]; src/symmetries.c:253-278 */
double oracle_prefactor_bUpdate(const int* irr, int updateCase, const int* sgs, int nrsy);

/* Column-major C_l = beta C_l + alpha op(A_l) op(B_l), l < L
 * (src/sparseblocks.c:133-158). trans: 0 = N, 1 = T. */
void oracle_dgemm_batched(int transa, int transb, int M, int N, int K, int L, double alpha,
                          const double* A, int lda, int64_t strideA, const double* B, int ldb,
                          int64_t strideB, double beta, double* C, int ldc, int64_t strideC);

/* Effective Hamiltonian (src/Heff.c). */
typedef struct oracle_heff oracle_heff;
int oracle_heff_init(oracle_heff** h, const t3ns_heff_desc* desc, const t3ns_rops_desc* ops,
                     const t3ns_site_desc* site);
int64_t oracle_heff_size(const oracle_heff* h);
/* result = Heff vec (result overwritten), OpenMP over output blocks. */
void oracle_heff_matvec(oracle_heff* h, const double* vec, double* result);
void oracle_heff_matvec_blocks(oracle_heff* h, const double* vec, double* result, int first,
                               int last, int stride);
void oracle_heff_block_stats(const oracle_heff* h, int64_t* npairs, int64_t* nentries,
                             int64_t* nterms, double* flops, double* block_flops);
void oracle_heff_diagonal(const oracle_heff* h, double* diag);
/* Plan table summary: number of (new,old) pairs, MPO entries and executed
 * GEMM chains, and the algorithmic flops of one matvec. */
void oracle_heff_plan_stats(const oracle_heff* h, int64_t* npairs, int64_t* nentries,
                            int64_t* nterms, double* flops);
/* The plan table itself, ordered by new block then pair then entry:
 * for each entry: newsb, oldsb, bestorder, mpo_group, sbops[3] (7 ints)
 * and the prefactor.  Returns the entry count; arrays may be NULL to count. */
int64_t oracle_heff_plan_table(const oracle_heff* h, int* ints, double* prefs);
void oracle_heff_destroy(oracle_heff* h);

/* rOperators_append_phys (src/rOperators_pUpdate.c:393-673 and
 * sum_unique_rOperators, src/rOperators_init.c:185-265, kick_zero_blocks,
 * src/sparseblocks.c:73-107) on the descriptor of include/t3ns_amd.h.
 * Output: for new operator k (k < max instr[2] + 1) its block offsets
 * (nblocks(hss_of_new[k]) + 1 entries, relative to the operator) appended
 * to `beginblock`, and its elements appended to `tel`.  Returns the total
 * element count; the arrays may be NULL to size them (the offsets then
 * need sum_k (nblocks_k + 1) entries). */
int64_t oracle_append_phys(const t3ns_append_desc* d, const t3ns_rops_desc* old,
                           int64_t* beginblock, double* tel);

/* update_rOperators_physical (src/rOperators_pUpdate.c:17-387) on the
 * descriptor of include/t3ns_amd.h; output as oracle_append_phys, one
 * operator per input operator. */
int64_t oracle_update_physical(const t3ns_pupdate_desc* d, const t3ns_rops_desc* pops,
                               int64_t* beginblock, double* tel);

/* update_rOperators_branching (src/rOperators_bUpdate.c:229-1101) on the
 * descriptor of include/t3ns_amd.h with ops[2]; output as
 * oracle_append_phys. */
int64_t oracle_update_branching(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops,
                                int64_t* beginblock, double* tel);
/* Forces contraction order `way` (0..5 of ways_to_update,
 * src/rOperators_bUpdate.c:813-820) for every block; -1 restores the
 * reference's choice of the cheapest. */
void oracle_set_bupdate_order(int way);

/* split_of_site (src/siteTensor_decompose.c:1403-1444) on the descriptor
 * of include/t3ns_amd.h; arrays of `out` are allocated here and released by
 * oracle_split_free.  Returns 0, or -1 when the site is not cut by one bond. */
int oracle_split_site(const t3ns_split_desc* d, t3ns_split_view* out);
void oracle_split_free(t3ns_split_view* v);

/* qr_step (src/siteTensor_decompose.c:1587-1630); arrays of `out` are
 * allocated here and released by oracle_qr_free.  Returns -1 when A and C
 * share no bond. */
int oracle_qr_step(const t3ns_qr_desc* d, t3ns_qr_view* out);
void oracle_qr_free(t3ns_qr_view* v);

/* makesiteTensor's contraction (src/siteTensor_init.c:609-800) on the
 * descriptor of include/t3ns_amd.h; tel: the multisite elements (zeroed
 * here first).  Returns 0, or -1 when an outer site is not bound to the
 * inner site. */
int oracle_make_site_tensor(const t3ns_makesite_desc* d, double* tel);

#ifdef __cplusplus
}
#endif

#endif
