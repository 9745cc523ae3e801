/*
 * ORACLE -- test infrastructure.  CPU restatement of the effective
 * Hamiltonian of the reference: src/Heff.c (init_Heffdata 1167-1221, the
 * qnB tables 108-484, first-run plan 695-914 and 1002-1047, replay 916-1000,
 * diagonal 1064-1108 and 1268-1312) and the merge-instruction grouping of
 * src/instructions.c:57-88.  Data comes in through the boundary descriptors
 * of include/t3ns_amd.h.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NS T3NS_MAX_SYMMETRIES
enum { T_NEW = 0, T_OLD = 1, T_OPS1 = 2, T_OPS2 = 3, T_OPS3 = 4, T_WORK1 = 5, T_WORK2 = 6 };

struct oss {
        int bond, n;
        int* irr;   /* n * NS */
        int* dims;
};

struct orops {
        int bond, is_left, P, nrhss, nrops;
        int* bbh;        /* nrhss + 1 */
        int64_t* qn;     /* nblocks * (P ? 3 : 1) */
        int* hss;        /* nrops */
        int64_t** bb;    /* per op: nblocks(op) + 1 */
        double** tel;    /* per op */
};

struct ontom {
        int oldsb, bestorder, nmbr;
        int (*sbops)[3];
        double* pref;
        int* MPO;
};

struct ocinfo {
        int tens[3];
        int trans[2];
        int M, N, K, L, lda, ldb, ldc;
        int64_t stride[3];
};

struct oracle_heff {
        int nrSyms, sgs[NS];
        struct oss mpo;
        int H;
        int* herm;
        unsigned char* singlet;
        struct oss symarr[4][3];
        int nrsites, isphys[4], nrblocks;
        int64_t* qn;
        int64_t* bb;
        struct orops ops[3];
        int rops_on_site[3];
        int step, nr_instr;
        int (*instr)[3];
        double* ipref;
        int nrMPOc;
        int* MPOc;
        int* MPOc_beg;
        int isdmrg, posB, map[3];
        int nr_qnB;
        int64_t* qnB;
        int* nr_to;
        int64_t** to;
        int** nrcombo;
        int*** combos;
        int** sbq;              /* blocks per qnB, -1 terminated */
        int (*dims)[3];
        int* nr_oldsb;
        struct ontom** ntom;
        int64_t worksize[2];
};

static const int ORD6[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
static const int ORD2[2][2] = {{0, 1}, {1, 0}};

static void copy_ss(struct oss* o, const t3ns_symsecs_desc* d)
{
        o->bond = d->bond;
        o->n = d->nrSecs;
        o->irr = malloc(sizeof(int) * (size_t) (d->nrSecs * NS + 1));
        o->dims = malloc(sizeof(int) * (size_t) (d->nrSecs + 1));
        memcpy(o->irr, d->irreps, sizeof(int) * (size_t) d->nrSecs * NS);
        memcpy(o->dims, d->dims, sizeof(int) * (size_t) d->nrSecs);
}

static void free_ss(struct oss* o)
{
        free(o->irr);
        free(o->dims);
}

static int nblocks_hss(const struct orops* r, int h) { return r->bbh[h + 1] - r->bbh[h]; }

static void copy_rops(struct orops* r, const t3ns_rops_desc* d)
{
        r->bond = d->bond;
        r->is_left = d->is_left;
        r->P = d->P_operator;
        r->nrhss = d->nrhss;
        r->nrops = d->nrops;
        r->bbh = malloc(sizeof(int) * (size_t) (d->nrhss + 1));
        memcpy(r->bbh, d->begin_blocks_of_hss, sizeof(int) * (size_t) (d->nrhss + 1));
        const int64_t nq = (int64_t) r->bbh[r->nrhss] * (r->P ? 3 : 1);
        r->qn = malloc(sizeof(int64_t) * (size_t) (nq + 1));
        memcpy(r->qn, d->qnumbers, sizeof(int64_t) * (size_t) nq);
        r->hss = malloc(sizeof(int) * (size_t) (d->nrops + 1));
        memcpy(r->hss, d->hss_of_ops, sizeof(int) * (size_t) d->nrops);
        r->bb = malloc(sizeof(int64_t*) * (size_t) (d->nrops + 1));
        r->tel = malloc(sizeof(double*) * (size_t) (d->nrops + 1));
        for (int o = 0; o < d->nrops; ++o) {
                const int nb = nblocks_hss(r, r->hss[o]);
                r->bb[o] = malloc(sizeof(int64_t) * (size_t) (nb + 1));
                for (int b = 0; b <= nb; ++b) r->bb[o][b] = d->beginblock[o][b] - d->beginblock[o][0];
                const int64_t n = r->bb[o][nb];
                r->tel[o] = malloc(sizeof(double) * (size_t) (n + 1));
                memcpy(r->tel[o], d->tel[o] + d->beginblock[o][0], sizeof(double) * (size_t) n);
        }
}

static void free_rops(struct orops* r)
{
        for (int o = 0; o < r->nrops; ++o) {
                free(r->bb[o]);
                free(r->tel[o]);
        }
        free(r->bb);
        free(r->tel);
        free(r->bbh);
        free(r->qn);
        free(r->hss);
}

/* ---------------------------------------------------- merge instructions */

/* src/instructions.c:57-88: stable sort of the instructions by their MPO
 * combination hss0 + hss1*H (+ hss2*H^2), grouped. */
static void merge_sort_instr(struct oracle_heff* h, const t3ns_heff_desc* d)
{
        int n = 0;
        for (int i = 0; i < d->merge_nr; ++i) {
                const int* in = &d->merge_instr[3 * i];
                if (in[0] >= 0 && in[1] >= 0 && in[2] >= 0) ++n;
        }
        int (*ins)[3] = malloc(sizeof(*ins) * (size_t) (n + 1));
        double* pr = malloc(sizeof(double) * (size_t) (n + 1));
        int* key = malloc(sizeof(int) * (size_t) (n + 1));
        int c = 0;
        for (int i = 0; i < d->merge_nr; ++i) {
                const int* in = &d->merge_instr[3 * i];
                if (in[0] < 0 || in[1] < 0 || in[2] < 0) continue;
                ins[c][0] = in[0];
                ins[c][1] = in[1];
                ins[c][2] = in[2];
                pr[c] = d->merge_pref[i];
                int t = 0;
                for (int j = d->merge_step - 1; j >= 0; --j)
                        t = t * h->H + h->ops[j].hss[in[j]];
                key[c] = t;
                ++c;
        }
        /* insertion sort: stable */
        int* idx = malloc(sizeof(int) * (size_t) (n + 1));
        for (int i = 0; i < n; ++i) {
                int j = i;
                while (j > 0 && key[idx[j - 1]] > key[i]) {
                        idx[j] = idx[j - 1];
                        --j;
                }
                idx[j] = i;
        }
        h->step = d->merge_step;
        h->nr_instr = n;
        h->instr = malloc(sizeof(*h->instr) * (size_t) (n + 1));
        h->ipref = malloc(sizeof(double) * (size_t) (n + 1));
        h->MPOc = malloc(sizeof(int) * (size_t) (n + 1));
        h->MPOc_beg = malloc(sizeof(int) * (size_t) (n + 2));
        h->nrMPOc = 0;
        for (int i = 0; i < n; ++i) {
                memcpy(h->instr[i], ins[idx[i]], sizeof(int) * 3);
                h->ipref[i] = pr[idx[i]];
                if (i == 0 || h->MPOc[h->nrMPOc - 1] != key[idx[i]]) {
                        h->MPOc_beg[h->nrMPOc] = i;
                        h->MPOc[h->nrMPOc++] = key[idx[i]];
                }
        }
        h->MPOc_beg[h->nrMPOc] = n;
        free(ins);
        free(pr);
        free(key);
        free(idx);
}

/* ------------------------------------------------------------- qnB tables */

static int cmp_i64(const void* a, const void* b)
{
        const int64_t x = *(const int64_t*) a, y = *(const int64_t*) b;
        return (x > y) - (x < y);
}

static int bsearch_i64(int64_t v, const int64_t* arr, int n)
{
        int lo = 0, hi = n - 1;
        while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                if (arr[mid] == v) return mid;
                if (arr[mid] < v) lo = mid + 1;
                else hi = mid - 1;
        }
        return -1;
}

/* (bra, ket) -> list of MPO sectors: qa[bra*dim + ket] = {count, h1, h2, ..}
 * src/Heff.c:118-174 */
static int** make_qnumbersarr(const struct orops* r, int dim)
{
        int** qa = calloc((size_t) dim * dim, sizeof(int*));
        int* cnt = calloc((size_t) dim * dim, sizeof(int));
        const int cpl = r->P ? 3 : 1;
        for (int pass = 0; pass < 2; ++pass) {
                int64_t prev = -1;
                int chss = 0;
                for (int i = 0; i < r->bbh[r->nrhss]; ++i) {
                        const int64_t cq = r->qn[(int64_t) cpl * i + cpl - 1];
                        while (i >= r->bbh[chss + 1]) ++chss;
                        if (prev == cq) continue;
                        const int bra = (int) (cq % dim);
                        const int ket = (int) ((cq / dim) % dim);
                        const int id = bra * dim + ket;
                        if (pass == 0) ++cnt[id];
                        else qa[id][++qa[id][0]] = chss;
                        prev = cq;
                }
                if (pass == 0) {
                        for (int i = 0; i < dim * dim; ++i) {
                                qa[i] = malloc(sizeof(int) * (size_t) (cnt[i] + 1));
                                qa[i][0] = 0;
                        }
                }
        }
        free(cnt);
        return qa;
}

static void free_qnumbersarr(int** qa, int dim)
{
        for (int i = 0; i < dim * dim; ++i) free(qa[i]);
        free(qa);
}

/* src/Heff.c:294-360 with find_next_old 242-280 and the MPO combinations
 * of count_or_make_MPOcombos 208-240 */
static void make_qnB_T3NS(struct oracle_heff* h)
{
        int dim[3];
        int** qa[3];
        for (int k = 0; k < 3; ++k) {
                dim[k] = h->symarr[h->posB][k].n;
                qa[k] = make_qnumbersarr(&h->ops[k], dim[k]);
        }
        const int64_t big = (int64_t) dim[0] * dim[1];
        const int H = h->H;
        for (int i = 0; i < h->nr_qnB; ++i) {
                int64_t q = h->qnB[i];
                const int n0 = (int) (q % dim[0]);
                q /= dim[0];
                const int n1 = (int) (q % dim[1]);
                const int n2 = (int) (q / dim[1]);
                h->to[i] = malloc(sizeof(int64_t) * (size_t) (h->nr_qnB + 1));
                h->nrcombo[i] = malloc(sizeof(int) * (size_t) (h->nr_qnB + 1));
                h->combos[i] = malloc(sizeof(int*) * (size_t) (h->nr_qnB + 1));
                int cnt = 0;
                for (int o1 = 0; o1 < dim[1]; ++o1) {
                        const int* l1 = qa[1][n1 * dim[1] + o1];
                        if (l1[0] == 0) continue;
                        for (int o0 = 0; o0 < dim[0]; ++o0) {
                                const int* l0 = qa[0][n0 * dim[0] + o0];
                                if (l0[0] == 0) continue;
                                const int64_t key = o0 + (int64_t) o1 * dim[0];
                                for (int loc = 0; loc < h->nr_qnB; ++loc) {
                                        if (h->qnB[loc] % big != key) continue;
                                        const int o2 = (int) (h->qnB[loc] / big);
                                        const int* l2 = qa[2][n2 * dim[2] + o2];
                                        int nc = 0;
                                        int* cl = malloc(sizeof(int) * (size_t) (l0[0] * l1[0] * l2[0] + 1));
                                        for (int a = 1; a <= l0[0]; ++a)
                                                for (int b = 1; b <= l1[0]; ++b)
                                                        for (int c = 1; c <= l2[0]; ++c) {
                                                                const int hc = h->herm[l2[c]];
                                                                if (h->singlet[l0[a] + (size_t) l1[b] * H + (size_t) hc * H * H])
                                                                        cl[nc++] = l0[a] + l1[b] * H + l2[c] * H * H;
                                                        }
                                        if (nc == 0) {
                                                free(cl);
                                                continue;
                                        }
                                        h->to[i][cnt] = h->qnB[loc];
                                        h->nrcombo[i][cnt] = nc;
                                        h->combos[i][cnt] = cl;
                                        ++cnt;
                                }
                        }
                }
                h->nr_to[i] = cnt;
        }
        for (int k = 0; k < 3; ++k) free_qnumbersarr(qa[k], dim[k]);
}

/* src/Heff.c:362-442 */
static void make_qnB_DMRG(struct oracle_heff* h)
{
        const int opid = h->ops[1].P;
        const struct orops* op = &h->ops[opid];
        const int N = op->bbh[op->nrhss];
        for (int i = 0; i < h->nr_qnB; ++i) {
                h->to[i] = malloc(sizeof(int64_t) * (size_t) (h->nr_qnB + 1));
                h->nrcombo[i] = calloc((size_t) h->nr_qnB + 1, sizeof(int));
                h->combos[i] = malloc(sizeof(int*) * (size_t) (h->nr_qnB + 1));
                int cnt = 0;
                int chss = 0;
                for (int j = 0; j < N; ++j) {
                        while (op->bbh[chss + 1] <= j) ++chss;
                        if (op->qn[3 * (int64_t) j] != h->qnB[i]) continue;
                        int k;
                        for (k = 0; k < cnt; ++k)
                                if (op->qn[3 * (int64_t) j + 1] == h->to[i][k]) break;
                        if (k == cnt) {
                                h->to[i][k] = op->qn[3 * (int64_t) j + 1];
                                h->combos[i][k] = malloc(sizeof(int) * (size_t) (op->nrhss + 1));
                                ++cnt;
                        }
                        const int c1 = opid == 1 ? h->herm[chss] : chss;
                        const int c2 = opid != 1 ? h->herm[chss] : chss;
                        h->combos[i][k][h->nrcombo[i][k]++] = c1 + c2 * h->H;
                }
                h->nr_to[i] = cnt;
        }
}

/* ------------------------------------------------------ first-run plan */

struct oidx {
        int64_t qn[4][2];
        int id[4][2][3];
        const int* irr[4][2][3];
        int dim[2][3];
};

/* src/Heff.c:695-739 */
static void fill_indexes(const struct oracle_heff* h, int sb, struct oidx* x, int tp)
{
        for (int i = 0; i < h->nrsites; ++i) {
                int64_t q = h->qn[(int64_t) h->nrsites * sb + i];
                x->qn[i][tp] = q;
                const struct oss* ss = h->symarr[i];
                x->id[i][tp][0] = (int) (q % ss[0].n);
                q /= ss[0].n;
                x->id[i][tp][1] = (int) (q % ss[1].n);
                x->id[i][tp][2] = (int) (q / ss[1].n);
                for (int k = 0; k < 3; ++k)
                        x->irr[i][tp][k] = &ss[k].irr[x->id[i][tp][k] * NS];
        }
        for (int i = 0; i < (h->isdmrg ? 2 : 3); ++i) {
                const int s = h->rops_on_site[i];
                if (!h->ops[i].P) {
                        const int loose = h->isdmrg ? 2 * (i == 1) : i;
                        x->dim[tp][i] = h->symarr[s][loose].dims[x->id[s][tp][loose]];
                } else {
                        const int loose = 2 * !h->ops[i].is_left;
                        x->dim[tp][i] = h->symarr[s][loose].dims[x->id[s][tp][loose]] *
                                h->symarr[s][1].dims[x->id[s][tp][1]];
                }
        }
        if (h->isdmrg) x->dim[tp][2] = 1;
}

/* src/Heff.c:507-524 */
static double calc_prefactor(const struct oracle_heff* h, const struct oidx* x, const int idMPO[3])
{
        double p = 1;
        int buf[2 * 3 * NS];
        for (int i = 0; i < (h->isdmrg ? 2 : 3); ++i) {
                if (!h->ops[i].P) continue;
                const int s = h->rops_on_site[i];
                for (int t = 0; t < 2; ++t)
                        for (int k = 0; k < 3; ++k)
                                memcpy(&buf[(t * 3 + k) * NS], x->irr[s][t][k], sizeof(int) * NS);
                p *= oracle_prefactor_add_P_operator(buf, h->ops[i].is_left, h->sgs, h->nrSyms);
        }
        for (int t = 0; t < 2; ++t)
                for (int k = 0; k < 3; ++k)
                        memcpy(&buf[(t * 3 + k) * NS], x->irr[h->posB][t][k], sizeof(int) * NS);
        int mpo[3 * NS];
        for (int k = 0; k < 3; ++k) {
                const int id = (h->isdmrg && k == 2) ? 0 : idMPO[k];
                memcpy(&mpo[k * NS], &h->mpo.irr[id * NS], sizeof(int) * NS);
        }
        p *= oracle_prefactor_combine_MPOs(buf, mpo, h->sgs, h->nrSyms, h->isdmrg,
                                           2 * !h->ops[1].P);
        return p;
}

static int cmp_qn3(const int64_t* a, const int64_t* b)
{
        for (int k = 2; k >= 0; --k)
                if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
        return 0;
}

/* src/Heff.c:760-796 */
static int find_operator_sb(const struct oracle_heff* h, const struct oidx* x, int i, int idMPO)
{
        const struct orops* r = &h->ops[i];
        const int s = h->rops_on_site[i];
        const int inner = r->P ? 2 * r->is_left : (h->isdmrg ? 2 * !r->is_left : i);
        const int din = h->symarr[s][inner].n;
        const int64_t qin = x->id[s][0][inner] + (int64_t) x->id[s][1][inner] * din +
                (int64_t) idMPO * din * din;
        const int n = nblocks_hss(r, idMPO);
        if (!r->P) return bsearch_i64(qin, r->qn + r->bbh[idMPO], n);
        const int64_t key[3] = {x->qn[s][0], x->qn[s][1], qin};
        const int64_t* arr = r->qn + 3 * (int64_t) r->bbh[idMPO];
        int lo = 0, hi = n - 1;
        while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                const int c = cmp_qn3(key, arr + 3 * (int64_t) mid);
                if (c == 0) return mid;
                if (c < 0) hi = mid - 1;
                else lo = mid + 1;
        }
        return -1;
}

/* src/Heff.c:526-583 */
static void prepare_cinfo_T3NS(const int dim[2][3], const int* map, struct ocinfo* ci, int ordernr)
{
        const int* order = ORD6[ordernr];
        int w[3] = {dim[1][0], dim[1][1], dim[1][2]};
        const int optype[3] = {T_OPS1, T_OPS2, T_OPS3};
        int res = T_OLD;
        for (int i = 0; i < 3; ++i) {
                const int o = order[i];
                ci[i].tens[0] = o == map[0] ? optype[o] : res;
                ci[i].tens[1] = o != map[0] ? optype[o] : res;
                ci[i].trans[0] = 0;
                ci[i].trans[1] = o == map[0] ? 0 : 1;
                w[o] = dim[0][o];
                ci[i].M = o == map[2] ? w[map[0]] * w[map[1]] : w[map[0]];
                ci[i].N = o == map[0] ? w[map[1]] * w[map[2]] : w[o];
                ci[i].K = dim[1][o];
                ci[i].L = o == map[1] ? w[map[2]] : 1;
                if (i != 2) {
                        ci[i].tens[2] = i == 0 ? T_WORK1 : T_WORK2;
                        res = ci[i].tens[2];
                } else {
                        ci[i].tens[2] = T_NEW;
                }
                ci[i].lda = ci[i].trans[0] == 0 ? ci[i].M : ci[i].K;
                ci[i].ldb = ci[i].trans[1] == 0 ? ci[i].K : ci[i].N;
                ci[i].ldc = ci[i].M;
                ci[i].stride[0] = (int64_t) ci[i].M * ci[i].K;
                ci[i].stride[1] = 0;
                ci[i].stride[2] = (int64_t) ci[i].M * ci[i].N;
        }
}

/* src/Heff.c:585-627 */
static void prepare_cinfo_DMRG(const int dim[2][3], struct ocinfo* ci, int ordernr)
{
        const int* order = ORD2[ordernr];
        int w[2] = {dim[1][0], dim[1][1]};
        int res = T_OLD;
        for (int i = 0; i < 2; ++i) {
                const int o = order[i];
                ci[i].tens[0] = o == 0 ? T_OPS1 : res;
                ci[i].tens[1] = o == 0 ? res : T_OPS2;
                ci[i].trans[0] = 0;
                ci[i].trans[1] = o == 0 ? 0 : 1;
                w[o] = dim[0][o];
                ci[i].M = w[0];
                ci[i].N = w[1];
                ci[i].K = dim[1][o];
                ci[i].L = 1;
                ci[i].tens[2] = i != 1 ? T_WORK1 : T_NEW;
                if (i != 1) res = T_WORK1;
                ci[i].lda = ci[i].trans[0] == 0 ? ci[i].M : ci[i].K;
                ci[i].ldb = ci[i].trans[1] == 0 ? ci[i].K : ci[i].N;
                ci[i].ldc = ci[i].M;
                ci[i].stride[0] = (int64_t) ci[i].M * ci[i].K;
                ci[i].stride[1] = 0;
                ci[i].stride[2] = (int64_t) ci[i].M * ci[i].N;
        }
}

/* src/Heff.c:629-666: cheapest contraction order */
static int best_order(const int dim[2][3], int isdmrg)
{
        int best = 0;
        long long nops = 0;
        for (int i = 0; i < (isdmrg ? 2 : 6); ++i) {
                long long cur = 0;
                long long w[3] = {dim[1][0], dim[1][1], dim[1][2]};
                for (int j = 0; j < (isdmrg ? 2 : 3); ++j) {
                        const int o = isdmrg ? ORD2[i][j] : ORD6[i][j];
                        cur += w[0] * w[1] * w[2] * dim[0][o];
                        w[o] = dim[0][o];
                }
                if (i == 0 || nops > cur) {
                        best = i;
                        nops = cur;
                }
        }
        return best;
}

static void do_contract(const struct ocinfo* c, double** tel, double alpha, double beta)
{
        oracle_dgemm_batched(c->trans[0], c->trans[1], c->M, c->N, c->K, c->L, alpha,
                             tel[c->tens[0]], c->lda, c->stride[0], tel[c->tens[1]], c->ldb,
                             c->stride[1], beta, tel[c->tens[2]], c->ldc, c->stride[2]);
}

/* operator element pointers for one instruction; 0 if a block is empty
 * (src/Heff.c:798-813) */
static int operator_tel(const struct oracle_heff* h, const int sb[3], const int* ins, double** tel)
{
        for (int i = 0; i < (h->isdmrg ? 2 : 3); ++i) {
                const struct orops* r = &h->ops[i];
                const int64_t* b = &r->bb[ins[i]][sb[i]];
                if (b[0] == b[1]) return 0;
                tel[i] = r->tel[ins[i]] + b[0];
        }
        return 1;
}

static void build_first_run(struct oracle_heff* h)
{
        const int n = h->nrblocks;
        h->dims = malloc(sizeof(*h->dims) * (size_t) (n + 1));
        h->nr_oldsb = calloc((size_t) n + 1, sizeof(int));
        h->ntom = calloc((size_t) n + 1, sizeof(struct ontom*));
        int64_t wsize[2] = {0, 0};
        for (int nq = 0; nq < h->nr_qnB; ++nq) {
                for (int* ps = h->sbq[nq]; *ps != -1; ++ps) {
                        const int newsb = *ps;
                        struct oidx x;
                        fill_indexes(h, newsb, &x, 0);
                        for (int k = 0; k < 3; ++k) h->dims[newsb][k] = x.dim[0][k];
                        struct ontom* list = malloc(sizeof(struct ontom) * (size_t) (n + 1));
                        int nl = 0;
                        for (int oq = 0; oq < h->nr_to[nq]; ++oq) {
                                const int oid = bsearch_i64(h->to[nq][oq], h->qnB, h->nr_qnB);
                                const int nc = h->nrcombo[nq][oq];
                                const int* mp = h->combos[nq][oq];
                                for (int* po = h->sbq[oid]; *po != -1; ++po) {
                                        fill_indexes(h, *po, &x, 1);
                                        struct ontom t = {.oldsb = *po, .nmbr = 0};
                                        t.bestorder = best_order((const int(*)[3]) x.dim, h->isdmrg);
                                        struct ocinfo ci[3];
                                        if (h->isdmrg) prepare_cinfo_DMRG((const int(*)[3]) x.dim, ci, t.bestorder);
                                        else prepare_cinfo_T3NS((const int(*)[3]) x.dim, h->map, ci, t.bestorder);
                                        int64_t c0 = (int64_t) ci[0].M * ci[0].N * ci[0].L;
                                        int64_t c1 = h->isdmrg ? 0 : (int64_t) ci[1].M * ci[1].N * ci[1].L;
                                        if (c0 > wsize[0]) wsize[0] = c0;
                                        if (c1 > wsize[1]) wsize[1] = c1;
                                        t.sbops = malloc(sizeof(*t.sbops) * (size_t) (nc + 1));
                                        t.pref = malloc(sizeof(double) * (size_t) (nc + 1));
                                        t.MPO = malloc(sizeof(int) * (size_t) (nc + 1));
                                        for (int c = 0; c < nc; ++c) {
                                                /* src/Heff.c:815-853 */
                                                const int g = mp[c];
                                                const int beg = h->MPOc_beg[g];
                                                if (h->MPOc_beg[g + 1] == beg) continue;
                                                int idMPO[3];
                                                idMPO[0] = h->ops[0].hss[h->instr[beg][0]];
                                                idMPO[1] = h->ops[1].hss[h->instr[beg][1]];
                                                idMPO[2] = h->isdmrg ? 0 : h->ops[2].hss[h->instr[beg][2]];
                                                const double pf = calc_prefactor(h, &x, idMPO);
                                                if (fabs(pf) < 1e-20) continue;
                                                int sb[3] = {-1, -1, -1};
                                                for (int k = 0; k < (h->isdmrg ? 2 : 3); ++k)
                                                        sb[k] = find_operator_sb(h, &x, k, idMPO[k]);
                                                if (sb[0] == -1 || sb[1] == -1) continue;
                                                t.MPO[t.nmbr] = g;
                                                t.pref[t.nmbr] = pf;
                                                memcpy(t.sbops[t.nmbr], sb, sizeof(sb));
                                                ++t.nmbr;
                                        }
                                        if (t.nmbr) list[nl++] = t;
                                        else {
                                                free(t.sbops);
                                                free(t.pref);
                                                free(t.MPO);
                                        }
                                }
                        }
                        h->ntom[newsb] = list;
                        h->nr_oldsb[newsb] = nl;
                }
        }
        h->worksize[0] = wsize[0] + 1;
        h->worksize[1] = wsize[1] + 1;
}

/* ------------------------------------------------------------ public API */

int oracle_heff_init(oracle_heff** out, const t3ns_heff_desc* d, const t3ns_rops_desc* ops,
                     const t3ns_site_desc* site)
{
        struct oracle_heff* h = calloc(1, sizeof(*h));
        h->nrSyms = d->nrSyms;
        memcpy(h->sgs, d->sgs, sizeof(h->sgs));
        copy_ss(&h->mpo, &d->mpo);
        h->H = d->mpo.nrSecs;
        h->herm = malloc(sizeof(int) * (size_t) h->H);
        memcpy(h->herm, d->hermitian, sizeof(int) * (size_t) h->H);
        h->singlet = malloc((size_t) h->H * h->H * h->H);
        memcpy(h->singlet, d->singlet, (size_t) h->H * h->H * h->H);
        h->nrsites = site->nrsites;
        h->nrblocks = site->nrblocks;
        memcpy(h->isphys, site->site_is_phys, sizeof(h->isphys));
        h->qn = malloc(sizeof(int64_t) * (size_t) (site->nrblocks * site->nrsites + 1));
        memcpy(h->qn, site->qnumbers, sizeof(int64_t) * (size_t) site->nrblocks * site->nrsites);
        h->bb = malloc(sizeof(int64_t) * (size_t) (site->nrblocks + 1));
        memcpy(h->bb, site->beginblock, sizeof(int64_t) * (size_t) (site->nrblocks + 1));
        for (int i = 0; i < site->nrsites; ++i)
                for (int k = 0; k < 3; ++k) copy_ss(&h->symarr[i][k], &d->symarr[i][k]);
        /* init_Heffdata, src/Heff.c:1167-1221 */
        h->isdmrg = (site->nrsites == 1 && site->site_is_phys[0]) ||
                (site->nrsites == 2 && site->site_is_phys[0] && site->site_is_phys[1]);
        for (int k = 0; k < (h->isdmrg ? 2 : 3); ++k) {
                copy_rops(&h->ops[k], &ops[k]);
                h->rops_on_site[k] = d->rops_on_site[k];
        }
        h->posB = 0;
        for (int i = 0; i < site->nrsites; ++i)
                if (!site->site_is_phys[i]) h->posB = i;
        if (h->isdmrg) h->posB = site->nrsites - 1;
        /* make_map, src/Heff.c:668-682 */
        const int sw = h->rops_on_site[1] != h->posB && h->rops_on_site[0] == h->posB;
        h->map[0] = sw ? 1 : 0;
        h->map[1] = sw ? 0 : 1;
        h->map[2] = 2;
        merge_sort_instr(h, d);
        /* unique qnB, src/Heff.c:444-470 */
        h->qnB = malloc(sizeof(int64_t) * (size_t) (h->nrblocks + 1));
        for (int i = 0; i < h->nrblocks; ++i) h->qnB[i] = h->qn[(int64_t) i * h->nrsites + h->posB];
        qsort(h->qnB, (size_t) h->nrblocks, sizeof(int64_t), cmp_i64);
        int u = 0;
        for (int i = 0; i < h->nrblocks; ++i)
                if (i == 0 || h->qnB[i] != h->qnB[u - 1]) h->qnB[u++] = h->qnB[i];
        h->nr_qnB = u;
        h->nr_to = calloc((size_t) u + 1, sizeof(int));
        h->to = calloc((size_t) u + 1, sizeof(int64_t*));
        h->nrcombo = calloc((size_t) u + 1, sizeof(int*));
        h->combos = calloc((size_t) u + 1, sizeof(int**));
        if (h->isdmrg) make_qnB_DMRG(h);
        else make_qnB_T3NS(h);
        /* blocks per qnB, src/Heff.c:1138-1165 */
        h->sbq = malloc(sizeof(int*) * (size_t) (u + 1));
        for (int i = 0; i < u; ++i) {
                int c = 0;
                h->sbq[i] = malloc(sizeof(int) * (size_t) (h->nrblocks + 1));
                for (int j = 0; j < h->nrblocks; ++j)
                        if (h->qn[(int64_t) j * h->nrsites + h->posB] == h->qnB[i]) h->sbq[i][c++] = j;
                h->sbq[i][c] = -1;
        }
        /* MPO combination -> group, src/Heff.c:1110-1136 (linear search) */
        for (int i = 0; i < u; ++i)
                for (int j = 0; j < h->nr_to[i]; ++j) {
                        int c = 0;
                        for (int k = 0; k < h->nrcombo[i][j]; ++k) {
                                int pos = -1;
                                for (int m = 0; m < h->nrMPOc; ++m)
                                        if (h->MPOc[m] == h->combos[i][j][k]) { pos = m; break; }
                                if (pos != -1) h->combos[i][j][c++] = pos;
                        }
                        h->nrcombo[i][j] = c;
                }
        build_first_run(h);
        *out = h;
        return 0;
}

int64_t oracle_heff_size(const oracle_heff* h) { return h->bb[h->nrblocks]; }

/* exec_secondrun / execute_heffcontr, src/Heff.c:916-1000 */
void oracle_heff_matvec(oracle_heff* h, const double* vec, double* result)
{
        oracle_heff_matvec_blocks(h, vec, result, 0, h->nrblocks, 1);
}

/* The same over output blocks first, first+stride, ... < last only (the
 * other entries of result are left untouched); used to time a bounded
 * sample of a matvec. */
void oracle_heff_matvec_blocks(oracle_heff* h, const double* vec, double* result, int first,
                               int last, int stride)
{
        for (int i = first; i < last; i += stride)
                for (int64_t e = h->bb[i]; e < h->bb[i + 1]; ++e) result[e] = 0;
        const int nsel = stride > 0 && last > first ? (last - first + stride - 1) / stride : 0;
#pragma omp parallel
        {
                double* tels[7];
                double* w1 = malloc(sizeof(double) * (size_t) h->worksize[0]);
                double* w2 = malloc(sizeof(double) * (size_t) h->worksize[1]);
                tels[T_WORK1] = w1;
                tels[T_WORK2] = w2;
#pragma omp for schedule(dynamic)
                for (int ii = 0; ii < nsel; ++ii) {
                        const int i = first + ii * stride;
                        int dims[2][3];
                        memcpy(dims[0], h->dims[i], sizeof(dims[0]));
                        tels[T_NEW] = result + h->bb[i];
                        for (int j = 0; j < h->nr_oldsb[i]; ++j) {
                                const struct ontom* t = &h->ntom[i][j];
                                memcpy(dims[1], h->dims[t->oldsb], sizeof(dims[1]));
                                struct ocinfo ci[3];
                                if (h->isdmrg) prepare_cinfo_DMRG((const int(*)[3]) dims, ci, t->bestorder);
                                else prepare_cinfo_T3NS((const int(*)[3]) dims, h->map, ci, t->bestorder);
                                tels[T_OLD] = (double*) vec + h->bb[t->oldsb];
                                for (int k = 0; k < t->nmbr; ++k) {
                                        const int g = t->MPO[k];
                                        for (int m = h->MPOc_beg[g]; m < h->MPOc_beg[g + 1]; ++m) {
                                                if (!operator_tel(h, t->sbops[k], h->instr[m], &tels[T_OPS1]))
                                                        continue;
                                                const double a = h->ipref[m] * t->pref[k];
                                                if (h->isdmrg) {
                                                        do_contract(&ci[0], tels, 1, 0);
                                                        do_contract(&ci[1], tels, a, 1);
                                                } else {
                                                        do_contract(&ci[0], tels, 1, 0);
                                                        do_contract(&ci[1], tels, 1, 0);
                                                        do_contract(&ci[2], tels, a, 1);
                                                }
                                        }
                                }
                        }
                }
                free(w1);
                free(w2);
        }
}

/* make_diagonal / diag_old_to_new_sb, src/Heff.c:1064-1108, 1268-1312 */
void oracle_heff_diagonal(const oracle_heff* h, double* diag)
{
        const int64_t size = h->bb[h->nrblocks];
        for (int64_t i = 0; i < size; ++i) diag[i] = 0;
#pragma omp parallel for schedule(dynamic)
        for (int nq = 0; nq < h->nr_qnB; ++nq) {
                int o;
                for (o = 0; o < h->nr_to[nq]; ++o)
                        if (h->to[nq][o] == h->qnB[nq]) break;
                if (o == h->nr_to[nq]) continue;
                for (int* ps = h->sbq[nq]; *ps != -1; ++ps) {
                        struct oidx x;
                        fill_indexes(h, *ps, &x, 0);
                        fill_indexes(h, *ps, &x, 1);
                        const int M = x.dim[1][h->map[0]];
                        const int N = x.dim[1][h->map[1]];
                        const int K = h->isdmrg ? 1 : x.dim[1][h->map[2]];
                        double* res = diag + h->bb[*ps];
                        for (int c = 0; c < h->nrcombo[nq][o]; ++c) {
                                const int g = h->combos[nq][o][c];
                                const int beg = h->MPOc_beg[g];
                                if (h->MPOc_beg[g + 1] == beg) continue;
                                int idMPO[3];
                                idMPO[0] = h->ops[0].hss[h->instr[beg][0]];
                                idMPO[1] = h->ops[1].hss[h->instr[beg][1]];
                                idMPO[2] = h->isdmrg ? 0 : h->ops[2].hss[h->instr[beg][2]];
                                int sb[3] = {-1, -1, -1};
                                for (int k = 0; k < (h->isdmrg ? 2 : 3); ++k)
                                        sb[k] = find_operator_sb(h, &x, k, idMPO[k]);
                                if (sb[0] < 0 || sb[1] < 0 || (!h->isdmrg && sb[2] < 0)) continue;
                                const double ps_ = calc_prefactor(h, &x, idMPO);
                                for (int m = beg; m < h->MPOc_beg[g + 1]; ++m) {
                                        double* t[3];
                                        if (!operator_tel(h, sb, h->instr[m], t)) continue;
                                        const double tp = h->ipref[m] * ps_;
                                        const double one = 1.0;
                                        for (int k = 0; k < K; ++k) {
                                                const double e3 = h->isdmrg ? one : t[h->map[2]][(int64_t) k * (K + 1)];
                                                for (int nn = 0; nn < N; ++nn) {
                                                        const double e2 = tp * e3 * t[h->map[1]][(int64_t) nn * (N + 1)];
                                                        for (int mm = 0; mm < M; ++mm)
                                                                res[mm + (int64_t) nn * M + (int64_t) k * M * N] +=
                                                                        e2 * t[h->map[0]][(int64_t) mm * (M + 1)];
                                                }
                                        }
                                }
                        }
                }
        }
}

void oracle_heff_plan_stats(const oracle_heff* h, int64_t* npairs, int64_t* nentries,
                            int64_t* nterms, double* flops)
{
        oracle_heff_block_stats(h, npairs, nentries, nterms, flops, NULL);
}

/* As oracle_heff_plan_stats, and the flops of every output block in
 * block_flops[nrblocks] when it is not NULL. */
void oracle_heff_block_stats(const oracle_heff* h, int64_t* npairs, int64_t* nentries,
                             int64_t* nterms, double* flops, double* block_flops)
{
        int64_t np = 0, ne = 0, nt = 0;
        double fl = 0;
        for (int i = 0; i < h->nrblocks; ++i) {
                const double fl_before = fl;
                for (int j = 0; j < h->nr_oldsb[i]; ++j) {
                        const struct ontom* t = &h->ntom[i][j];
                        ++np;
                        int dims[2][3];
                        memcpy(dims[0], h->dims[i], sizeof(dims[0]));
                        memcpy(dims[1], h->dims[t->oldsb], sizeof(dims[1]));
                        struct ocinfo ci[3];
                        if (h->isdmrg) prepare_cinfo_DMRG((const int(*)[3]) dims, ci, t->bestorder);
                        else prepare_cinfo_T3NS((const int(*)[3]) dims, h->map, ci, t->bestorder);
                        double f = 0;
                        for (int s = 0; s < (h->isdmrg ? 2 : 3); ++s)
                                f += 2.0 * ci[s].M * (double) ci[s].N * ci[s].K * ci[s].L;
                        for (int k = 0; k < t->nmbr; ++k) {
                                ++ne;
                                const int g = t->MPO[k];
                                for (int m = h->MPOc_beg[g]; m < h->MPOc_beg[g + 1]; ++m) {
                                        double* tt[3];
                                        if (!operator_tel(h, t->sbops[k], h->instr[m], tt)) continue;
                                        ++nt;
                                        fl += f;
                                }
                        }
                }
                if (block_flops) block_flops[i] = fl - fl_before;
        }
        if (npairs) *npairs = np;
        if (nentries) *nentries = ne;
        if (nterms) *nterms = nt;
        if (flops) *flops = fl;
}

int64_t oracle_heff_plan_table(const oracle_heff* h, int* ints, double* prefs)
{
        int64_t c = 0;
        for (int i = 0; i < h->nrblocks; ++i)
                for (int j = 0; j < h->nr_oldsb[i]; ++j) {
                        const struct ontom* t = &h->ntom[i][j];
                        for (int k = 0; k < t->nmbr; ++k, ++c) {
                                if (ints) {
                                        int* r = ints + 7 * c;
                                        r[0] = i;
                                        r[1] = t->oldsb;
                                        r[2] = t->bestorder;
                                        r[3] = t->MPO[k];
                                        r[4] = t->sbops[k][0];
                                        r[5] = t->sbops[k][1];
                                        r[6] = h->isdmrg ? -1 : t->sbops[k][2];
                                }
                                if (prefs) prefs[c] = t->pref[k];
                        }
                }
        return c;
}

void oracle_heff_destroy(oracle_heff* h)
{
        if (!h) return;
        free_ss(&h->mpo);
        free(h->herm);
        free(h->singlet);
        for (int i = 0; i < h->nrsites; ++i)
                for (int k = 0; k < 3; ++k) free_ss(&h->symarr[i][k]);
        for (int k = 0; k < (h->isdmrg ? 2 : 3); ++k) free_rops(&h->ops[k]);
        free(h->qn);
        free(h->bb);
        free(h->instr);
        free(h->ipref);
        free(h->MPOc);
        free(h->MPOc_beg);
        for (int i = 0; i < h->nr_qnB; ++i) {
                for (int j = 0; j < h->nr_to[i]; ++j) free(h->combos[i][j]);
                free(h->to[i]);
                free(h->nrcombo[i]);
                free(h->combos[i]);
                free(h->sbq[i]);
        }
        free(h->nr_to);
        free(h->to);
        free(h->nrcombo);
        free(h->combos);
        free(h->sbq);
        for (int i = 0; i < h->nrblocks; ++i) {
                for (int j = 0; j < h->nr_oldsb[i]; ++j) {
                        free(h->ntom[i][j].sbops);
                        free(h->ntom[i][j].pref);
                        free(h->ntom[i][j].MPO);
                }
                free(h->ntom[i]);
        }
        free(h->ntom);
        free(h->nr_oldsb);
        free(h->dims);
        free(h->qnB);
        free(h);
}
