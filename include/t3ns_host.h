/*
 * t3ns_amd host layer: solve contexts (the reference's globals made
 * explicit) and step objects that produce the inputs of the Heff boundary in
 * include/t3ns_amd.h.  These replace the calling pattern of the reference's
 * tests (tests/test1.c.in:31-54: make_network, readinteraction,
 * preparebookkeeper) and the per-step glue of optimize_siteTensor
 * (src/optimize_network.c:188-226).
 */
#ifndef T3NS_HOST_H
#define T3NS_HOST_H

#include <stdint.h>

#include "t3ns_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct t3ns_ctx t3ns_ctx;
typedef struct t3ns_step t3ns_step;

/* Solve context.  netw_text: contents of a .netw network file
 * (src/network.c:207-354).  model: "NN_HUBBARD" with params {t, U}
 * (src/hamiltonian_nn_hubbard.c:75-84).  sgs/ts: symmetry groups and target
 * irreps (enum symmetrygroup order, include/symmetries.h:34); Z2 is added
 * when absent (src/bookkeeper.c:495-513).  max_dim and minocc as in
 * preparebookkeeper(NULL, max_dim, 1, minocc, NULL) (src/bookkeeper.c:725). */
int t3ns_ctx_create(t3ns_ctx** c, const char* netw_text, const char* model, const double* params,
                    int nparams, const int* sgs, const int* ts, int nsym, int max_dim, int minocc);
/* Quantum-chemistry Hamiltonian from FCIDUMP text (src/hamiltonian_qc.c:152,
 * QC_make_hamiltonian; header and integrals as read at src/hamiltonian_qc.c:783-900).
 * sgs must hold U1 and U1 or SU2, optionally one point group whose ORBSYM
 * labels the FCIDUMP carries. */
int t3ns_ctx_create_qc(t3ns_ctx** c, const char* netw_text, const char* fcidump_text,
                       const int* sgs, const int* ts, int nsym, int max_dim, int minocc);
/* The optimisation steps of one sweep as next_opt_step yields them
 * (src/network.c:529-727) with at most maxsites sites per step: for step i,
 * sites_out[4*i .. 4*i+3] (unused slots -1) and nsites_out[i].  Returns the
 * number of steps (entries beyond cap are counted, not written). */
int t3ns_ctx_sweep_steps(t3ns_ctx* c, int maxsites, int* sites_out, int* nsites_out, int cap);
int t3ns_ctx_destroy(t3ns_ctx* c);
/* nr_bonds, sites, psites of the network; nrSyms and sgs after Z2 insertion. */
int t3ns_ctx_info(const t3ns_ctx* c, int* nr_bonds, int* sites, int* psites, int* nrsyms,
                  int* sgs);
/* The two sites that bond `bond` joins (netw.bonds, -1 at a border). */
int t3ns_ctx_bond_sites(const t3ns_ctx* c, int bond, int* sites2);
/* The three bonds of site `site` (get_bonds_of_site, src/network.c:430) and
 * whether it is a physical site. */
int t3ns_ctx_site_bonds(const t3ns_ctx* c, int site, int* bonds3, int* is_phys);
/* on = 0: the problem builders (t3ns_step_random, t3ns_step_*_desc,
 * t3ns_ctx_*_desc except t3ns_ctx_site_tensor) give structure only, with
 * zero elements, for callers that put their own data in; on = 1 (the
 * default): seeded N(0,1) elements as documented at each builder. */
int t3ns_ctx_fill_seeded(t3ns_ctx* c, int on);
/* Replaces the sectors of bond `bond` (irreps [nrsecs][T3NS_MAX_SYMMETRIES],
 * dims [nrsecs]), as a decomposition leaves them: the bookkeeper update of
 * adapt_UV_tensors_and_kick_empties (src/siteTensor_decompose.c:1360-1401). */
int t3ns_ctx_set_bond(t3ns_ctx* c, int bond, int nrsecs, const int* irreps, const int* dims);
/* A one-site tensor of site `site` over the context's current sectors
 * (init_1siteTensor, src/siteTensor_init.c:64-108) with seeded N(0,1) elements;
 * the arrays live as long as the context and the elements are writable. */
int t3ns_ctx_site_tensor(t3ns_ctx* c, int site, uint64_t seed, int* nrblocks, const int64_t** qnumbers,
                         const int64_t** beginblock, double** tel);
/* Sector count and dims of bond `bond` (reference numbering). */
int t3ns_ctx_bond(const t3ns_ctx* c, int bond, int* nrsecs, int* totaldims, int* dims_out,
                  int dims_cap);

/* A multisite optimisation object over `sites` (2 physical sites, or a
 * branching site with up to three physical neighbours) whose renormalised
 * operators hold deterministic pseudo-random blocks (seed): the Heff inputs
 * of one step without a preceding sweep.  Used by parity tests. */
int t3ns_step_random(t3ns_step** s, t3ns_ctx* c, const int* sites, int nsites, uint64_t seed);
/* The rOperators_append_phys problem behind physical leg `leg` of a step:
 * the operators of the bond on the far side of the appended site (N(0,1)
 * entries from `seed`, or the unit operator at a border) and the append
 * descriptor.  Pointers stay valid until the step is destroyed. */
int t3ns_step_append_desc(t3ns_step* s, int leg, uint64_t seed, t3ns_append_desc* d,
                          t3ns_rops_desc* old);
/* An update_rOperators_physical problem for physical leg `leg` of a step:
 * the bond's sectors are replaced by a truncated set (a seeded random
 * subset of the internal sectors with dims 0..maxdim, as a decomposition
 * leaves them; maxdim 0: all sectors at the full dimension of the product
 * space, no truncation), the one-site tensor of the physical site gets
 * N(0,1) entries, and the step's physical operators of that leg are the
 * input.  The context is left as it was.  The arrays live in the step
 * until the next call for the same leg with another seed or maxdim. */
int t3ns_step_pupdate_desc(t3ns_step* s, int leg, uint64_t seed, int maxdim, t3ns_pupdate_desc* d,
                           t3ns_rops_desc* pops);
/* A makesiteTensor problem for a step: seeded N(0,1) one-site tensors of
 * the step's sites over the sectors their bonds had before the multisite
 * structure replaced the internal ones, and the step's multisite
 * structure.  The arrays live in the step until the next call with another
 * seed. */
int t3ns_step_makesite_desc(t3ns_step* s, uint64_t seed, t3ns_makesite_desc* d);
/* An update_rOperators_branching problem at branching site `site` for
 * new operators on its bond `update_case`: seeded N(0,1) old operators on
 * the two other bonds (vacuum operators at a border), a seeded site tensor,
 * the bUpdate instructions and the new operators' structure.  The arrays
 * live as long as the context. */
int t3ns_ctx_bupdate_desc(t3ns_ctx* c, int site, int update_case, uint64_t seed,
                          t3ns_bupdate_desc* d, t3ns_rops_desc* ops /* [2] */);
/* The rOperators_append_phys problem on bond `bond` seen from side
 * `is_left` (the physical site on that side is appended): seeded N(0,1)
 * operators on the site's far bond (the unit operator at a border) as
 * `old`, which a caller may replace by its own operators of that bond, and
 * the append descriptor: the append of init_rops
 * (src/optimize_network.c:436-451) and of the Heff legs next to a physical
 * site (src/optimize_network.c:164-176).  The arrays live as long as the
 * context. */
int t3ns_ctx_append_desc(t3ns_ctx* c, int bond, int is_left, uint64_t seed, t3ns_append_desc* d,
                         t3ns_rops_desc* old);
/* The update_rOperators_physical problem on bond `bond` seen from side
 * `is_left`: a seeded N(0,1) one-site tensor of the physical site on that
 * side over the context's own sectors (no truncation) and the structure of
 * the new operators; the physical operators are the caller's (the result
 * of the append on the same bond).  The arrays live as long as the
 * context. */
int t3ns_ctx_pupdate_desc(t3ns_ctx* c, int bond, int is_left, uint64_t seed, t3ns_pupdate_desc* d);
/* A split of the step's multisite tensor (its N(0,1) elements normalised):
 * site index `siteV` is split off with the given truncation settings. */
int t3ns_step_split_desc(t3ns_step* s, int siteV, int minD, int maxD, double truncerr,
                         char truncType, t3ns_split_desc* d);
/* A qr_step problem: seeded N(0,1) one-site tensors of sites siteA and
 * siteC (neighbours) over the context's sectors.  The arrays live as long
 * as the context. */
int t3ns_ctx_qr_desc(t3ns_ctx* c, int siteA, int siteC, uint64_t seed, t3ns_qr_desc* d);
int t3ns_step_destroy(t3ns_step* s);
int64_t t3ns_step_size(const t3ns_step* s);
/* Fills the Heff boundary descriptors; pointers stay valid while s lives. */
int t3ns_step_heff_desc(t3ns_step* s, t3ns_heff_desc* desc, t3ns_rops_desc* ops,
                        t3ns_site_desc* site);

/* Checkpoints in the reference's HDF5 layout (T3NScalc.h5,
 * src/io_to_disk.c:446-487): /network, /bookkeeper (the context's current
 * bond sectors), /hamiltonian and /T3NS (one single-site tensor per network
 * site: nrblocks[i] blocks, qnumbers[i][nrblocks[i]], beginblock[i]
 * [nrblocks[i] + 1], tel[i][beginblock[i][nrblocks[i]]]); /rOps holds nrOps
 * and no operators (a restored calculation rebuilds them from the
 * tensors).  HDF5 is loaded at run time (libhdf5.so.103, or the path in
 * T3NS_HDF5_LIB).  Replaces write_to_disk (src/io_to_disk.c:446-466). */
int t3ns_ckpt_write(const t3ns_ctx* c, const char* path, int nsites, const int* nrblocks,
                    const int64_t* const* qnumbers, const int64_t* const* beginblock,
                    const double* const* tel);

/* Reads a checkpoint (written here or by the reference: its rOperator
 * groups are counted, not read).  Replaces read_from_disk
 * (src/io_to_disk.c:468-487) with an object in place of the globals. */
typedef struct t3ns_ckpt t3ns_ckpt;
int t3ns_ckpt_open(t3ns_ckpt** k, const char* path);
int t3ns_ckpt_info(const t3ns_ckpt* k, int* sites, int* nr_bonds, int* ham_type,
                   int* nr_rops_groups);
/* Site tensor `site`: pointers valid until t3ns_ckpt_close. */
int t3ns_ckpt_site(const t3ns_ckpt* k, int site, int* nrsites, int* nrblocks,
                   const int64_t** qnumbers, const int64_t** beginblock, const double** tel);
/* A new context with the checkpoint's network, symmetries, Hamiltonian and
 * bond sectors. */
int t3ns_ckpt_context(const t3ns_ckpt* k, t3ns_ctx** c);
int t3ns_ckpt_close(t3ns_ckpt* k);

#ifdef __cplusplus
}
#endif

#endif /* T3NS_HOST_H */
