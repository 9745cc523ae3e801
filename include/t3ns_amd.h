/*
 * t3ns_amd: MI355X-native effective-Hamiltonian matvec, Davidson and
 * operator renormalisation for block-sparse three-legged tree tensor
 * networks.  C ABI: plain pointers and sizes, no C++ or torch types.
 *
 * Drop-in boundary.  The reference's Heff path is
 *   init_Heffdata(struct Heffdata*, const struct rOperators*, const struct siteTensor*)
 *                                                  include/Heff.h:166-167
 *   matvecT3NS(const double* vec, double* result, void* vdata)   include/Heff.h:177
 *   make_diagonal(const struct Heffdata*)                        include/Heff.h:185
 *   destroy_Heffdata(struct Heffdata*)                           include/Heff.h:192
 *   davidson(...)                                                include/davidson.h:56-60
 * and reads the process globals bookie / netw / ham.  A shim in the
 * reference flattens its structs and globals into the descriptors below and
 * calls t3ns_heff_init / t3ns_heff_matvec / t3ns_heff_diagonal /
 * t3ns_heff_destroy / t3ns_davidson.
 *
 * All functions return 0 on success and a negative code on error; the
 * message of the last error of the calling thread is t3ns_last_error().
 */
#ifndef T3NS_AMD_H
#define T3NS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T3NS_MAX_SYMMETRIES 5

enum t3ns_status {
    T3NS_OK = 0,
    T3NS_ERR_INVALID = -1,   /* bad argument or inconsistent input */
    T3NS_ERR_DEVICE = -2,    /* HIP runtime error or no device */
    T3NS_ERR_NOMEM = -3,     /* device allocation failed */
    T3NS_ERR_SOLVER = -4     /* eigensolver failure (dsyev in the reference) */
};

enum t3ns_precision {
    T3NS_FP64 = 0, /* plain path: FP64 storage and arithmetic */
    T3NS_FP32 = 1, /* FP32 storage and arithmetic for the bulk matvec */
    T3NS_BF16 = 2  /* BF16 operator storage, each operator block scaled by a
                      power of two; FP32 matrix-core arithmetic and
                      accumulation, FP32 stage intermediates */
};

/* Symmetry sectors of one bond: struct symsecs, include/symsecs.h:32-54. */
typedef struct {
    int bond;             /* reference bond number (-1 for the MPO bond) */
    int nrSecs;
    const int* irreps;    /* [nrSecs][T3NS_MAX_SYMMETRIES] */
    const int* dims;      /* [nrSecs] */
} t3ns_symsecs_desc;

/* Renormalised operators: struct rOperators, include/rOperators.h:91-123.
 * Offsets are 64-bit element offsets (the reference uses int). */
typedef struct {
    int bond;
    int is_left;
    int P_operator;
    int nrhss;
    const int* begin_blocks_of_hss;   /* [nrhss+1] */
    const int64_t* qnumbers;          /* [nblocks * (P_operator ? 3 : 1)] */
    int nrops;
    const int* hss_of_ops;            /* [nrops] */
    const int64_t* const* beginblock; /* [nrops] arrays of nblocks(op)+1 offsets */
    const double* const* tel;         /* [nrops] host element arrays */
} t3ns_rops_desc;

/* The multisite tensor being optimised: struct siteTensor,
 * include/siteTensor.h:73-95 (elements are passed to the matvec). */
typedef struct {
    int nrsites;
    int sites[4];
    int site_is_phys[4];
    int nrblocks;
    const int64_t* qnumbers;   /* [nrblocks * nrsites] */
    const int64_t* beginblock; /* [nrblocks+1] */
} t3ns_site_desc;

/* Everything init_Heffdata reads from the globals bookie / netw / ham and
 * from fetch_merge (src/Heff.c:1167-1221, src/instructions.c:188-222). */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];        /* enum symmetrygroup values */
    t3ns_symsecs_desc mpo;               /* MPO bond sectors */
    const int* hermitian;                /* [mpo.nrSecs]: hermitian_symsec */
    const unsigned char* singlet;        /* [H^3]: MPO_couples_to_singlet({a,b,c}) at a+b*H+c*H*H */
    t3ns_symsecs_desc symarr[4][3];      /* sectors of the three bonds of each site */
    int rops_on_site[3];                 /* site index each operator attaches to */
    int merge_step;                      /* 2 (DMRG) or 3 (T3NS) */
    int merge_nr;                        /* raw merge instructions as the model emits them */
    const int* merge_instr;              /* [merge_nr * 3] */
    const double* merge_pref;            /* [merge_nr] */
} t3ns_heff_desc;

typedef struct t3ns_heff t3ns_heff;

const char* t3ns_last_error(void);

/* Number of visible devices (0 if none). */
int t3ns_device_count(void);

/* init_Heffdata (include/Heff.h:166): builds the dispatch plan on the host,
 * uploads plan tables and operator blocks to HBM.  ops[2] is ignored for a
 * two-site (DMRG) object.  precision: enum t3ns_precision. */
int t3ns_heff_init(t3ns_heff** h, const t3ns_heff_desc* desc, const t3ns_rops_desc* ops,
                   const t3ns_site_desc* site, int precision);

/* An FP32 object on an FP64 one's plan and operator arenas (the arenas
 * copied and converted on the device, the plan copied; only its FP32 work
 * lists are built on the host), for the low-precision operator of
 * t3ns_davidson_refined: no operator data crosses the host link.  No
 * counterpart in the reference (its Heff is FP64 only).  precision must be
 * T3NS_FP32; the result is independent of `base` afterwards (either may be
 * destroyed first); autotune and set_tuning are refused on it. */
int t3ns_heff_derive(t3ns_heff** h, const t3ns_heff* base, int precision);

/* matvecT3NS (include/Heff.h:177): result = Heff vec, host FP64 arrays of
 * t3ns_heff_size() elements; result is overwritten. */
int t3ns_heff_matvec(t3ns_heff* h, const double* vec, double* result);

/* Same on device-resident FP64 arrays, enqueued on `stream` (hipStream_t,
 * may be NULL for the default stream). */
int t3ns_heff_matvec_device(t3ns_heff* h, const double* dvec, double* dresult, void* stream);

/* on != 0: matvecs on a stream other than the default ones replay a HIP
 * graph captured at the first call with the same (input, output) device
 * arrays, one launch instead of the matvec's sequence of kernels (the
 * Davidson of this library reuses its search-vector slots, so its matvecs
 * replay after the first max_vecs); on = 0 (default) drops them. */
int t3ns_heff_use_graphs(t3ns_heff* h, int on);

/* make_diagonal (include/Heff.h:185): diag(Heff) into a host array. */
int t3ns_heff_diagonal(t3ns_heff* h, double* diag);

/* destroy_Heffdata (include/Heff.h:192). */
int t3ns_heff_destroy(t3ns_heff* h);

/* Plan statistics: vector length, work-list sizes and the algorithmic flops
 * of one matvec (sum of 2*M*N*K*L over the GEMMs the reference executes). */
int64_t t3ns_heff_size(const t3ns_heff* h);
int t3ns_heff_stats(const t3ns_heff* h, int64_t* nblocks, int64_t* npairs, int64_t* nterms,
                    double* flops);

/* Batched matvecs (no counterpart in the reference: its solves run one at
 * a time): the matvecs of n objects of one precision go through one launch
 * per kernel, their work units sharing the persistent grid, so that the
 * matvecs of concurrent small solves fill the GPU together.  Each result is
 * bitwise that of the object's own matvec.  A batch object keeps its
 * buffers between calls; calls on one batch must use one stream.
 * batch_matvec takes device arrays (dx[i], dy[i] of t3ns_heff_size(hs[i])
 * doubles, on `stream`); batch_matvec_host host arrays (each object at
 * most once), synchronously. */
typedef struct t3ns_heff_batch t3ns_heff_batch;
int t3ns_heff_batch_create(t3ns_heff_batch** b);
int t3ns_heff_batch_destroy(t3ns_heff_batch* b);
int t3ns_heff_batch_matvec(t3ns_heff_batch* b, t3ns_heff* const* hs, const double* const* dx,
                           double* const* dy, int n, void* stream);
int t3ns_heff_batch_matvec_host(t3ns_heff_batch* b, t3ns_heff* const* hs, const double* const* x,
                                double* const* y, int n);

/* Autotuner (no counterpart in the reference: its matvec has no tunable
 * work decomposition).  Times matvecs under at most max_trials settings of
 * the work-list cost model (grid shape, fused-operand price, chunk price,
 * fused middle stage on/off), keeps the fastest and records it for the
 * object's block-size histogram; an object whose histogram was tuned before
 * in this process takes the recorded setting without timing.  out[0..n):
 * chunk price, fused-operand price, fused (0/1), schedule, ms per matvec
 * under the result, ms under the setting before tuning (0 when taken from
 * the record), settings timed, taken from the record (0/1), workgroups per
 * CU.  Results change by rounding only. */
int t3ns_heff_autotune(t3ns_heff* h, int max_trials, double* out, int n);
/* The work-list setting t[0..5) read or imposed: chunk price, fused-operand
 * price, fused (0/1), schedule (0 cost, 1 block, 2 xcd, 3 xcdblock),
 * workgroups per CU (1, 2 or 4: 8, 4 or 2 waves each with the whole, half
 * or quarter LDS pool; 0 picks by the object's structure). */
int t3ns_heff_get_tuning(const t3ns_heff* h, double* t);
int t3ns_heff_set_tuning(t3ns_heff* h, const double* t);
/* Forgets every recorded tuning of this process. */
void t3ns_tune_cache_clear(void);

/* Work-list statistics of the device kernels, out[0..n): chunks, runs,
 * pairs whose input block is read from global memory, pairs whose stage
 * outputs use global scratch, matrix-core steps per matvec (total and per
 * stage A/B/C), global scratch elements per workgroup, work units, output
 * blocks split over several units, pairs whose middle stage is fused into
 * the last, workgroups per CU of the grid.  Returns the number of
 * statistics available. */
int t3ns_heff_kernel_stats(const t3ns_heff* h, double* out, int n);

/* The dispatch table the reference builds in its first matvec
 * (struct secondrun, include/Heff.h:36-51; src/Heff.c:815-914), ordered by
 * output block, then pair, then MPO entry.  Per entry 7 ints (newsb, oldsb,
 * bestorder, mpo_group, sbops[0..2]; sbops[2] = -1 for DMRG) and the symmetry
 * prefactor.  Returns the number of entries; arrays may be NULL to count. */
int64_t t3ns_heff_plan_table(const t3ns_heff* h, int* ints, double* prefs);

/* The same table built on the host only (no device needed), plus the
 * algorithmic flops of one matvec.  Returns the entry count or a negative
 * status. */
int64_t t3ns_plan_table(const t3ns_heff_desc* desc, const t3ns_rops_desc* ops,
                        const t3ns_site_desc* site, int* ints, double* prefs, double* flops);

/* Lowest eigenpair of the effective Hamiltonian: davidson(result, energy,
 * size, max_vecs, keep_deflate, davidson_tol, max_its, diagonal, matvecT3NS,
 * vdata) of include/davidson.h:56-60 with the Heff handle as the operator,
 * everything resident on the device.  result: host guess in, eigenvector
 * out; diagonal: host diagonal or NULL (computed).  Returns 0 converged,
 * 1 stopped at max_its (src/davidson.c:327), T3NS_ERR_SOLVER if the
 * projected eigenproblem failed (-1 in the reference), other negatives on
 * error. */
int t3ns_davidson(t3ns_heff* h, double* result, double* energy, int max_vecs, int keep_deflate,
                  double tol, int max_its, const double* diagonal, int* its);

/* The same eigenpair with a reduced-precision operator for the bulk of the
 * iterations.  The reduced operator's relative error e is measured on the
 * guess (one matvec of each); Davidson with h_low's matvec (e.g.
 * T3NS_FP32) runs until the residual is below max(tol, switch_tol) or
 * below 2 e |theta| (where its eigenvector is as close to the FP64 one as
 * e allows).  If that met tol, one FP64 matvec gives the FP64 Rayleigh
 * quotient and residual, the result when that residual is below tol
 * (its64 = 0); otherwise the same search space (its keep_deflate lowest
 * Ritz vectors, their images recomputed with h64) continues with FP64
 * matvecs down to tol.  The energy is an FP64 Rayleigh quotient.  Same
 * return codes. */
int t3ns_davidson_refined(t3ns_heff* h_low, t3ns_heff* h64, double* result, double* energy,
                          int max_vecs, int keep_deflate, double tol, double switch_tol,
                          int max_its, const double* diagonal, int* its_low, int* its64);

/* Davidson on any symmetric operator given as a device matvec callback
 * (y = A x on device pointers, on `stream`); dx: device guess in,
 * eigenvector out; ddiag: device diagonal. */
typedef void (*t3ns_dev_matvec)(const double* x, double* y, void* stream, void* vdata);
int t3ns_davidson_dev(int64_t size, double* dx, const double* ddiag, t3ns_dev_matvec matvec,
                      void* vdata, void* stream, double* energy, int max_vecs, int keep_deflate,
                      double tol, int max_its, int* its);

/* t3ns_davidson_dev without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc/hipFree and no allocating library
 * routine; the solver's buffers (search vectors, images, reductions) live
 * in a fixed 64 MiB device array per GPU, held until the process ends, so
 * (2 max_vecs + 1 + keep_deflate) * size doubles (plus small reduction
 * buffers) must fit in it: about 133 000 elements at max_vecs = 30; larger
 * sizes and max_vecs above 64 return T3NS_ERR_INVALID.  Every call on a GPU
 * shares that GPU's array and leaves its contents behind, so calls that
 * overlap on one GPU race.  The callback must not allocate either for the
 * whole solve to be allocation-free (t3ns_heff_matvec_device does not). */
int t3ns_davidson_dev_noalloc(int64_t size, double* dx, const double* ddiag,
                              t3ns_dev_matvec matvec, void* vdata, void* stream, double* energy,
                              int max_vecs, int keep_deflate, double tol, int max_its, int* its);

/* ------------------------------------------------------------------------
 * Reference-shaped entry points: the signatures the reference's caller
 * optimize_siteTensor (src/optimize_network.c:188-226) uses, so a shim maps
 * its calls one to one (init_Heffdata -> t3ns_heff_init, destroy_Heffdata ->
 * t3ns_heff_destroy; the rest below).
 * ---------------------------------------------------------------------- */

/* matvecT3NS(vec, result, vdata) of include/Heff.h:177 (src/Heff.c:1049):
 * vdata is the t3ns_heff handle; host arrays of t3ns_heff_size() doubles,
 * result overwritten.  void like the reference: a failure prints the
 * message and exits (EXIT_FAILURE), as the reference's fatal errors do
 * (src/macros.c:49-57). */
void t3ns_matvecT3NS(const double* vec, double* result, void* vdata);

/* The same operator as a t3ns_dev_matvec callback (device arrays, stream),
 * vdata the t3ns_heff handle: t3ns_davidson_dev(size, dx, ddiag,
 * t3ns_heff_dev_matvec, h, ...) runs a solve with no host round trips.
 * Failures print and exit like t3ns_matvecT3NS. */
void t3ns_heff_dev_matvec(const double* dx, double* dy, void* stream, void* vdata);

/* make_diagonal (include/Heff.h:185, src/Heff.c:1268-1312): a malloc'ed
 * array of t3ns_heff_size() doubles that the caller frees with free();
 * NULL on error (message in t3ns_last_error). */
double* t3ns_make_diagonal(t3ns_heff* h);

/* davidson (include/davidson.h:56-60, src/davidson.c:259-328) with the
 * reference's arguments: host guess in / eigenvector out, host diagonal,
 * and the caller's host matvec callback (e.g. t3ns_matvecT3NS with the
 * t3ns_heff handle as vdat).  The search space stays on the device; each
 * matvec copies one vector to the host and its image back.  Returns 0
 * converged, 1 stopped at max_its, -1 on failure (the reference's dsyev
 * failure, or a bad argument or device error: message in t3ns_last_error). */
int t3ns_davidson_cb(double* result, double* energy, int size, int max_vecs, int keep_deflate,
                     double davidson_tol, int max_its, const double* diagonal,
                     void (*matvec)(const double*, double*, void*), void* vdat);

/* sparse_eigensolve (include/wrapper_solvers.h:56-60,
 * src/wrapper_solvers.c:107-128): solver "D" is t3ns_davidson_cb; any
 * other solver (PRIMME) returns -1. */
int t3ns_sparse_eigensolve(double* result, double* energy, int size, int max_vecs,
                           int keep_deflate, double tol, int max_its, const double* diagonal,
                           void (*matvec)(const double*, double*, void*), void* vdat,
                           const char solver[]);

/* ------------------------------------------------------------------------
 * Renormalisation: appending a physical site to renormalised operators.
 * rOperators_append_phys(nr, or) (src/rOperators_pUpdate.c:665-673 with
 * unique_rOperators_ap 642-663, pAppend_block 537-640, sum_unique_rOperators
 * src/rOperators_init.c:228-265): the operators of the bond on the far side
 * of a physical site become operators acting on (bond x site), P_operator = 1.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];
    int is_left;                     /* of the old and the new operators */
    t3ns_symsecs_desc site_ss[3];    /* sectors of the appended site's bonds (alpha, i, beta) */
    t3ns_symsecs_desc new_ss[3];     /* bra, ket and MPO sectors of the new operators' bond */
    t3ns_symsecs_desc old_ss[3];     /* bra, ket and MPO sectors of the old operators' bond */
    t3ns_symsecs_desc mpo;           /* MPO sectors */
    /* pUpdate instructions of the old bond as fetch_pUpdate sorts them
     * (src/instructions.c:116-140): (old op, site op, new op), prefactor,
     * MPO sector of every new operator */
    int nr_instr;
    const int* instr;                /* [nr_instr * 3] */
    const double* pref;              /* [nr_instr] */
    const int* hss_of_new;           /* [number of new operators] */
    /* site operators: MPO sector and matrix elements el_siteop(op, bra, ket) */
    int nr_siteops;
    const int* siteop_hss;           /* [nr_siteops] */
    int physdim;                     /* sectors of the physical bond */
    const double* siteop_el;         /* [nr_siteops * physdim * physdim], (op*physdim + bra)*physdim + ket */
    /* tprods_ham(new MPO sector, site): pairs (site-op sector, old-op sector) */
    const int* tprod_begin;          /* [mpo.nrSecs + 1] */
    const int* tprods;               /* [tprod_begin[H] * 2] */
    /* structure of the new operators before empty blocks are dropped */
    int nrhss;
    const int* begin_blocks_of_hss;  /* [nrhss + 1] */
    const int64_t* qnumbers;         /* [nblocks * 3] */
    const int64_t* const* hss_beginblock; /* per MPO sector: block offsets [nblocks(hss) + 1] */
} t3ns_append_desc;

/* Renormalised operators owned by the library (host copy). */
typedef struct t3ns_rops t3ns_rops;

/* Work of one device execution (renormalisation, site-tensor operations). */
typedef struct {
    double flops;      /* algorithmic flops of the device kernels */
    double bytes;      /* inputs read once + output written once */
    double kernel_ms;  /* elapsed time of the device kernels */
} t3ns_work;


/* rOperators_append_phys on the device: *out receives the new operators
 * (all-zero blocks dropped as kick_zero_blocks does, src/sparseblocks.c:73-107). */
int t3ns_append_phys(const t3ns_append_desc* desc, const t3ns_rops_desc* old, t3ns_rops** out);
/* Descriptor view of library-owned operators (valid until destroyed). */
int t3ns_rops_view(const t3ns_rops* r, t3ns_rops_desc* d);
int t3ns_rops_destroy(t3ns_rops* r);

/* ------------------------------------------------------------------------
 * Renormalisation: update_rOperators_physical (src/rOperators_pUpdate.c:
 * 364-387 with make_usb_to_osb 114-138, select_site_blocks 210-271,
 * get_upd_aide 273-321, pUpdate_block 323-362): physical operators on an
 * internal bond become operators on the new (truncated) bond,
 *   left:  O' = T_herm^T O T,   right: O' = T_herm O T^T,
 * with T the optimised one-site tensor and the adjoint/pUpdate prefactors.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];
    int is_left;
    t3ns_symsecs_desc site_ss[3];    /* sectors of the tensor's bonds (the updated bond truncated) */
    t3ns_symsecs_desc internal_ss;   /* sectors the physical operators were built on */
    t3ns_symsecs_desc mpo;
    /* the one-site tensor */
    int T_nrblocks;
    const int64_t* T_qnumbers;       /* [T_nrblocks] over site_ss */
    const int64_t* T_beginblock;     /* [T_nrblocks + 1] */
    const double* T_tel;
    /* structure of the updated operators (one quantum number per block over
     * (bra, ket, MPO) of the updated bond), before empty blocks are dropped */
    int nrhss;
    const int* begin_blocks_of_hss;
    const int64_t* qnumbers;
    const int64_t* const* hss_beginblock;
} t3ns_pupdate_desc;

int t3ns_update_physical(const t3ns_pupdate_desc* desc, const t3ns_rops_desc* pops,
                         t3ns_rops** out);

/* Work of the device execution that produced `r`: algorithmic flops, bytes
 * (inputs read once, output written once) and the kernels' elapsed time.
 * No counterpart in the reference (it reports wall-clock timers only,
 * src/optimize_network.c:255-279). */
int t3ns_rops_work(const t3ns_rops* r, double* flops, double* bytes, double* kernel_ms);

/* ------------------------------------------------------------------------
 * Renormalisation: update_rOperators_branching (src/rOperators_bUpdate.c:
 * 229-248 with update_unique_ops_T3NS 637-677, update_newblock_w_MPO_set
 * 746-785, update_selected_blocks 1043-1072, prefactor_bUpdate
 * src/symmetries.c:253-278, then sum_unique_rOperators
 * src/rOperators_init.c:228-265): operators on one bond of a branching site
 * from the operators on its two other bonds and the site tensor T,
 *   new = sum pref * T_adj . (O0 x O1) . T   per bUpdate instruction.
 * ops[0] sits on bond (update_case == 0 ? 1 : 0) and is a left operator;
 * ops[1] on bond (update_case == 2 ? 1 : 2), left only for update_case 2.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];
    int update_case;                 /* bond of the site carrying the new operators */
    t3ns_symsecs_desc site_ss[3];    /* the site's bonds (bra and ket sectors alike) */
    t3ns_symsecs_desc mpo;
    int T_nrblocks;
    const int64_t* T_qnumbers;       /* over site_ss, sorted */
    const int64_t* T_beginblock;
    const double* T_tel;
    /* bUpdate instructions, sorted as the reference sorts them
     * (src/instructions.c:152-186) */
    int nr_instr;
    const int* instr;                /* [nr_instr * 3]: op of ops[0], op of ops[1], new op */
    const double* pref;
    const int* hss_of_new;
    /* per MPO sector h of a new operator, the pairs (sector of ops[0],
     * sector of ops[1]) tprod[tprod_begin[h] .. tprod_begin[h+1]) that make
     * it (tprods_ham, src/hamiltonian.c:227-251) */
    const int* tprod_begin;          /* [nrhss + 1] */
    const int* tprods;               /* [2 * tprod_begin[nrhss]] */
    /* structure of the new operators (all blocks present) */
    int nrhss;
    const int* begin_blocks_of_hss;
    const int64_t* qnumbers;
    const int64_t* const* hss_beginblock;
} t3ns_bupdate_desc;

int t3ns_update_branching(const t3ns_bupdate_desc* desc, const t3ns_rops_desc* ops /* [2] */,
                          t3ns_rops** out);

/* The host half of t3ns_update_branching alone (no device): out[0..n) =
 * seconds taking in the descriptor (operator and tensor copies), seconds
 * planning, block products planned, unique-operator blocks written.  For
 * sizing and timing the planner. */
int t3ns_bupdate_plan_stats(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops, double* out,
                            int n);

/* ------------------------------------------------------------------------
 * decompose_siteTensor, one split (src/siteTensor_decompose.c:1403-1444,
 * split_of_site, with init_svddata 587-887, svdblocks 1050-1072, selectS
 * 1196-1289, SVD_copy_from_mem 989-1048, adapt_UV_tensors_and_kick_empties
 * 1360-1401, norm_tensor src/siteTensor_misc.c:244-250): the site `siteV`
 * of a multisite tensor A is split off by a truncated SVD per sector of the
 * bond it shares with the rest.  U (the remaining sites, singular values
 * multiplied in, normalised) and V (the site) come back with the cut bond's
 * kept sectors.  Repeating the split gives the reference's HOSVD
 * (1533-1585).  The SVDs run on the device (one-sided Jacobi); singular
 * vectors agree with LAPACK's up to the sign of each pair.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];
    int nrsites;                         /* 2..4 */
    int siteV;                           /* index of the site split off */
    t3ns_symsecs_desc ss[4][3];          /* sectors of every site's bonds (bond ids set) */
    int nrblocks;
    const int64_t* qnumbers;             /* [nrblocks * nrsites] */
    const int64_t* beginblock;           /* [nrblocks + 1] */
    const double* tel;                   /* A, normalised */
    int minD, maxD;                      /* struct SvalSelect, include/siteTensor.h:425-431 */
    double truncerr;
    char truncType;                      /* 'E' (entropy) or 'N' (norm) */
} t3ns_split_desc;

typedef struct t3ns_split t3ns_split;
typedef struct {
    int U_nrsites, U_nrblocks;
    const int64_t* U_qnumbers;           /* [U_nrblocks * U_nrsites] */
    const int64_t* U_beginblock;
    const double* U_tel;
    int V_nrblocks;
    const int64_t* V_qnumbers;
    const int64_t* V_beginblock;
    const double* V_tel;
    int bond_nrSecs;                     /* the cut bond's kept sectors */
    const int* bond_irreps;              /* [bond_nrSecs][T3NS_MAX_SYMMETRIES] */
    const int* bond_dims;
    int nrsectors;                       /* sectors before truncation */
    const int* nsing;                    /* [nrsectors] singular values per sector */
    const int* kept;                     /* [nrsectors] kept per sector */
    const double* sing;                  /* all singular values, sector after sector, descending */
    double norm[2], entropy[2];          /* before / after truncation (struct SelectRes) */
    double norm_U;                       /* the norm divided out of U */
    int sweeps;                          /* Jacobi sweeps */
} t3ns_split_view;

int t3ns_split_site(const t3ns_split_desc* desc, t3ns_split** out, t3ns_work* work);
int t3ns_split_get(const t3ns_split* s, t3ns_split_view* view);
int t3ns_split_destroy(t3ns_split* s);

/* ------------------------------------------------------------------------
 * decompose_siteTensor for a one-site object: qr_step (src/siteTensor_
 * decompose.c:1587-1630 with qr 313-355, qrblocks 270-311, multiplyR
 * 416-490, R_svd 550-574).  A is QR-decomposed over the bond it shares with
 * the next orthogonality centre C; Q replaces A (the bond's dim per sector
 * becomes min(rows, dim)) and R is multiplied into C.  Householder QR on the
 * device with LAPACK's sign convention; the entropy of R's spectrum is
 * reported as the reference does.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrSyms;
    int sgs[T3NS_MAX_SYMMETRIES];
    t3ns_symsecs_desc ssA[3], ssC[3];    /* bonds of A and C (bond ids set) */
    int A_nrblocks;
    const int64_t* A_qnumbers;
    const int64_t* A_beginblock;
    const double* A_tel;
    int C_nrblocks;
    const int64_t* C_qnumbers;
    const int64_t* C_beginblock;
    const double* C_tel;
} t3ns_qr_desc;

typedef struct t3ns_qr t3ns_qr;
typedef struct {
    int nrblocks_Q;                      /* blocks and quantum numbers of A */
    const int64_t* Q_beginblock;
    const double* Q_tel;
    int nrblocks_B;                      /* blocks and quantum numbers of C */
    const int64_t* B_beginblock;
    const double* B_tel;
    int nrsectors;                       /* sectors of the shared bond */
    const int* dims;                     /* its new dim per sector */
    const int* nsing;
    const double* sing;                  /* R's singular values, sector after sector */
    double entropy;
} t3ns_qr_view;

int t3ns_qr_step(const t3ns_qr_desc* desc, t3ns_qr** out, t3ns_work* work);
int t3ns_qr_get(const t3ns_qr* q, t3ns_qr_view* view);
int t3ns_qr_destroy(t3ns_qr* q);

/* ------------------------------------------------------------------------
 * makesiteTensor (src/siteTensor_init.c:785-800 with init_makeinfo 609-649,
 * init_contractinfo 651-760, contractsiteTensors 768-783): the multisite
 * tensor's elements from two to four one-site tensors, contracted over the
 * internal bonds (the bonds the sites share, recognised by their ids in
 * `new_ss`).  Multisite blocks with a missing or empty one-site block stay
 * zero.  The one-site case of the reference is a pointer copy, not a
 * computation, and is not offered here.
 * ---------------------------------------------------------------------- */
typedef struct {
    int nrsites;                         /* 2..4 */
    int innersite;                       /* the branching site; the second of two physical sites */
    t3ns_symsecs_desc new_ss[4][3];      /* multisite sectors (internal bonds at unit dims) */
    t3ns_symsecs_desc old_ss[4][3];      /* sectors of the one-site tensors */
    int nrblocks;                        /* multisite blocks */
    const int64_t* qnumbers;             /* [nrblocks * nrsites] over new_ss */
    const int64_t* beginblock;           /* [nrblocks + 1] */
    int site_nrblocks[4];
    const int64_t* site_qnumbers[4];     /* [site_nrblocks[i]] over old_ss[i], sorted */
    const int64_t* site_beginblock[4];   /* [site_nrblocks[i] + 1] */
    const double* site_tel[4];
} t3ns_makesite_desc;

/* tel: [beginblock[nrblocks]] doubles, overwritten.  work may be NULL. */
int t3ns_make_site_tensor(const t3ns_makesite_desc* desc, double* tel, t3ns_work* work);

/* Leg dimensions of every block of the multisite tensor in operator-leg
 * order (dims[3*b + leg]; leg 2 is 1 for two-site DMRG objects), as
 * fill_indexes derives them (src/Heff.c:695-739).  Returns the block count
 * or a negative status. */
int t3ns_plan_block_dims(const t3ns_heff_desc* desc, const t3ns_rops_desc* ops,
                         const t3ns_site_desc* site, int* dims);

#ifdef __cplusplus
}
#endif

#endif /* T3NS_AMD_H */
