/* ORACLE -- test infrastructure, not product code.
 *
 * Plain-C restatement of appending a physical site to renormalised
 * operators, following the reference step by step:
 *   unique operators from the sorted pUpdate instructions
 *       (src/rOperators_pUpdate.c:393-426, src/instructions.c:224-252),
 *   pAppend_block for every block of the unique operators
 *       (src/rOperators_pUpdate.c:537-640),
 *   sum_unique_rOperators (src/rOperators_init.c:185-265),
 *   kick_zero_blocks (src/sparseblocks.c:73-107).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void indexize3(int* id, int64_t qn, const t3ns_symsecs_desc* ss)
{
        const int64_t n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
        id[0] = (int) (qn % n0);
        id[1] = (int) ((qn / n0) % n1);
        id[2] = (int) (qn / (n0 * n1));
}

static int64_t qntypize3(const int* id, const t3ns_symsecs_desc* ss)
{
        return id[0] + (int64_t) id[1] * ss[0].nrSecs + (int64_t) id[2] * ss[0].nrSecs * ss[1].nrSecs;
}

static int bsearch_qn(int64_t v, const int64_t* a, int n)
{
        int lo = 0, hi = n - 1;
        while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                if (a[mid] == v) return mid;
                if (a[mid] < v) lo = mid + 1;
                else hi = mid - 1;
        }
        return -1;
}

static const int* irrep_of(const t3ns_symsecs_desc* s, int i)
{
        return s->irreps + (size_t) i * T3NS_MAX_SYMMETRIES;
}

/* get_next_unique_instr (src/instructions.c:224-252) */
static int next_unique(int* cur, const t3ns_append_desc* d)
{
        if (*cur == -1) {
                ++*cur;
                return d->nr_instr > 0;
        }
        const int* o = d->instr + 3 * *cur;
        const int ho = d->hss_of_new[o[2]];
        for (++*cur; *cur < d->nr_instr; ++*cur) {
                const int* n = d->instr + 3 * *cur;
                if (o[0] != n[0] || o[1] != n[1] || ho != d->hss_of_new[n[2]]) return 1;
        }
        return 0;
}

int64_t oracle_append_phys(const t3ns_append_desc* d, const t3ns_rops_desc* old,
                           int64_t* out_bb, double* out_tel)
{
        const int H = d->mpo.nrSecs;
        const int pbond = 2 * !d->is_left;
        /* unique operators */
        int nu = 0;
        int cur = -1;
        while (next_unique(&cur, d)) ++nu;
        int* u_po = malloc(sizeof(int) * (nu + 1));
        int* u_so = malloc(sizeof(int) * (nu + 1));
        int* u_hss = malloc(sizeof(int) * (nu + 1));
        cur = -1;
        for (int u = 0; next_unique(&cur, d); ++u) {
                u_po[u] = d->instr[3 * cur + 0];
                u_so[u] = d->instr[3 * cur + 1];
                u_hss[u] = d->hss_of_new[d->instr[3 * cur + 2]];
        }
        /* unique operator storage: one calloc'ed arena per operator */
        double** utel = malloc(sizeof(double*) * (nu + 1));
        for (int u = 0; u < nu; ++u) {
                const int h = u_hss[u];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                utel[u] = calloc((size_t) d->hss_beginblock[h][nb] + 1, sizeof(double));
        }
        /* pAppend_block for every new block */
        const int nblocks = d->begin_blocks_of_hss[d->nrhss];
        for (int usb = 0; usb < nblocks; ++usb) {
                int ids[3][3];
                indexize3(ids[0], d->qnumbers[3 * usb + 0], d->site_ss);
                indexize3(ids[1], d->qnumbers[3 * usb + 1], d->site_ss);
                indexize3(ids[2], d->qnumbers[3 * usb + 2], d->new_ss);
                const int hssn = ids[2][2];
                ids[2][pbond == 0 ? 2 : 0] = hssn;
                const int ublock = usb - d->begin_blocks_of_hss[hssn];
                for (int p = d->tprod_begin[hssn]; p < d->tprod_begin[hssn + 1]; ++p) {
                        const int hsss = d->tprods[2 * p];
                        const int hsso = d->tprods[2 * p + 1];
                        ids[2][pbond] = hsso;
                        ids[2][1] = hsss;
                        const int id_old[3] = {ids[0][pbond], ids[1][pbond], ids[2][pbond]};
                        const int64_t oqn = qntypize3(id_old, d->old_ss);
                        int irr[3][3][T3NS_MAX_SYMMETRIES];
                        for (int i = 0; i < 2; ++i)
                                for (int j = 0; j < 3; ++j)
                                        memcpy(irr[i][j], irrep_of(&d->site_ss[j], ids[i][j]), sizeof(irr[i][j]));
                        for (int j = 0; j < 3; ++j)
                                memcpy(irr[2][j], irrep_of(&d->mpo, ids[2][j]), sizeof(irr[2][j]));
                        const double pref = oracle_prefactor_pAppend(&irr[0][0][0], d->is_left,
                                                                     d->sgs, d->nrSyms);
                        const int onb = old->begin_blocks_of_hss[hsso + 1] - old->begin_blocks_of_hss[hsso];
                        const int oblock = bsearch_qn(oqn, old->qnumbers + old->begin_blocks_of_hss[hsso], onb);
                        if (oblock == -1 || fabs(pref) < 1e-20) continue;
                        for (int u = 0; u < nu; ++u) {
                                const int po = u_po[u], so = u_so[u];
                                const int chss[3] = {old->hss_of_ops[po], d->siteop_hss[so], u_hss[u]};
                                if (!(chss[0] == hsso && chss[1] == hsss && chss[2] == hssn)) continue;
                                const int64_t* obb = old->beginblock[po];
                                const int64_t N = obb[oblock + 1] - obb[oblock];
                                const double site_el = pref * d->siteop_el[((size_t) so * d->physdim + ids[0][1]) *
                                                                           d->physdim + ids[1][1]];
                                const double* o = old->tel[po] + obb[oblock];
                                double* t = utel[u] + d->hss_beginblock[hssn][ublock];
                                for (int64_t j = 0; j < N; ++j) t[j] = site_el * o[j];
                        }
                }
        }
        /* sum_unique_rOperators */
        int nops = 0;
        for (int i = 0; i < d->nr_instr; ++i)
                if (d->instr[3 * i + 2] + 1 > nops) nops = d->instr[3 * i + 2] + 1;
        double** ntel = malloc(sizeof(double*) * (nops + 1));
        for (int k = 0; k < nops; ++k) {
                const int h = d->hss_of_new[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                ntel[k] = calloc((size_t) d->hss_beginblock[h][nb] + 1, sizeof(double));
        }
        int u = 0;
        for (int i = 0; i < d->nr_instr; ++i) {
                const int* in = d->instr + 3 * i;
                if (i != 0) {
                        const int* pi = d->instr + 3 * (i - 1);
                        if (in[0] != pi[0] || in[1] != pi[1] ||
                            d->hss_of_new[in[2]] != d->hss_of_new[pi[2]])
                                ++u;
                }
                const int h = d->hss_of_new[in[2]];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                const int64_t N = d->hss_beginblock[h][nb];
                for (int64_t j = 0; j < N; ++j) ntel[in[2]][j] += d->pref[i] * utel[u][j];
        }
        /* kick_zero_blocks and output */
        int64_t total = 0, nbb = 0;
        for (int k = 0; k < nops; ++k) {
                const int h = d->hss_of_new[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                const int64_t* bb = d->hss_beginblock[h];
                int64_t pos = 0;
                if (out_bb) out_bb[nbb] = 0;
                for (int b = 0; b < nb; ++b) {
                        int nonzero = 0;
                        for (int64_t j = bb[b]; j < bb[b + 1]; ++j)
                                if (!(fabs(ntel[k][j]) < 1e-20)) { nonzero = 1; break; }
                        const int64_t n = nonzero ? bb[b + 1] - bb[b] : 0;
                        if (out_tel && n) memcpy(out_tel + total + pos, ntel[k] + bb[b], sizeof(double) * n);
                        pos += n;
                        if (out_bb) out_bb[nbb + b + 1] = pos;
                }
                nbb += nb + 1;
                total += pos;
        }
        for (int i = 0; i < nu; ++i) free(utel[i]);
        for (int k = 0; k < nops; ++k) free(ntel[k]);
        free(utel);
        free(ntel);
        free(u_po);
        free(u_so);
        free(u_hss);
        return total;
}

/* ------------------------------------------------------------------------
 * update_rOperators_physical (src/rOperators_pUpdate.c:364-387):
 * make_usb_to_osb (114-138), select_site_blocks (210-271), the dgemm order
 * of get_upd_aide (273-321), pUpdate_block (323-362), cleanup_update's
 * kick_zero_blocks (179-188).
 * ---------------------------------------------------------------------- */

static int search_irrep(const int* irrep, const t3ns_symsecs_desc* ss, int nrsy)
{
        for (int i = 0; i < ss->nrSecs; ++i) {
                int k = 0;
                for (; k < nrsy; ++k)
                        if (irrep_of(ss, i)[k] != irrep[k]) break;
                if (k == nrsy) return i;
        }
        return -1;
}

/* Th^T (left) or Th (right) as an N0 x M0 matrix */
static double adj_el(const double* th, int il, int N0, int M0, int n, int m)
{
        return il ? th[m + (int64_t) n * M0] : th[n + (int64_t) m * N0];
}

/* T (left) or T^T (right) as an M1 x N1 matrix */
static double site_el(const double* t, int il, int M1, int N1, int m, int n)
{
        return il ? t[m + (int64_t) n * M1] : t[n + (int64_t) m * N1];
}

int64_t oracle_update_physical(const t3ns_pupdate_desc* d, const t3ns_rops_desc* po,
                               int64_t* out_bb, double* out_tel)
{
        const int il = d->is_left;
        const int ns = d->nrSyms;
        const int ub = il ? 2 : 0;
        t3ns_symsecs_desc uss[3] = {d->site_ss[ub], d->site_ss[ub], d->mpo};
        t3ns_symsecs_desc oss0[3] = {d->site_ss[0], d->site_ss[1], d->site_ss[2]};
        oss0[ub] = d->internal_ss;
        t3ns_symsecs_desc oss2[3] = {d->internal_ss, d->internal_ss, d->mpo};

        const int N = d->begin_blocks_of_hss[d->nrhss];
        const int NN = po->begin_blocks_of_hss[po->nrhss];
        /* make_usb_to_osb; entries the reference leaves unset (no old block
         * above the new one) are taken as NN */
        int* usb_to_osb = malloc(sizeof(int) * (N + 1));
        for (int i = 0; i <= N; ++i) usb_to_osb[i] = NN;
        usb_to_osb[0] = 0;
        int osb = 0;
        for (int usb = 0; usb < N; ++usb) {
                int id[3], nid[3];
                indexize3(id, d->qnumbers[usb], uss);
                for (int k = 0; k < 2; ++k) nid[k] = search_irrep(irrep_of(&uss[k], id[k]), &oss2[k], ns);
                nid[2] = id[2];
                const int64_t qn = qntypize3(nid, oss2);
                for (; osb < NN; ++osb)
                        if (qn < po->qnumbers[3 * osb + 2]) {
                                usb_to_osb[usb + 1] = osb;
                                break;
                        }
        }

        double** ntel = malloc(sizeof(double*) * (po->nrops + 1));
        for (int k = 0; k < po->nrops; ++k) {
                const int h = po->hss_of_ops[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                ntel[k] = calloc((size_t) d->hss_beginblock[h][nb] + 1, sizeof(double));
        }

        for (int usb = 0; usb < N; ++usb) {
                int newhss = 0;
                while (usb >= d->begin_blocks_of_hss[newhss + 1]) ++newhss;
                const int usb2 = usb - d->begin_blocks_of_hss[newhss];
                for (int o = usb_to_osb[usb]; o < usb_to_osb[usb + 1]; ++o) {
                        const int osb2 = o - po->begin_blocks_of_hss[newhss];
                        /* select_site_blocks */
                        const int64_t* cqn = po->qnumbers + 3 * o;
                        int ids[3][3];
                        indexize3(ids[0], cqn[0], oss0);
                        indexize3(ids[1], cqn[1], oss0);
                        indexize3(ids[2], cqn[2], oss2);
                        int Tids[2][3], bad = 0;
                        for (int k = 0; k < 3; ++k) {
                                Tids[0][k] = search_irrep(irrep_of(&oss0[k], ids[1][k]), &d->site_ss[k], ns);
                                Tids[1][k] = search_irrep(irrep_of(&oss0[k], ids[0][k]), &d->site_ss[k], ns);
                                bad |= Tids[0][k] < 0 || Tids[1][k] < 0;
                        }
                        if (bad) continue;
                        const int tsb = bsearch_qn(qntypize3(Tids[0], d->site_ss), d->T_qnumbers, d->T_nrblocks);
                        const int thsb = bsearch_qn(qntypize3(Tids[1], d->site_ss), d->T_qnumbers, d->T_nrblocks);
                        if (tsb < 0 || thsb < 0) continue;
                        if (d->T_beginblock[tsb + 1] == d->T_beginblock[tsb] ||
                            d->T_beginblock[thsb + 1] == d->T_beginblock[thsb])
                                continue;
                        const double* T = d->T_tel + d->T_beginblock[tsb];
                        const double* Th = d->T_tel + d->T_beginblock[thsb];
                        const int N0 = d->site_ss[2 * il].dims[Tids[1][2 * il]];
                        const int N1 = d->site_ss[2 * il].dims[Tids[0][2 * il]];
                        const int M0 = d->site_ss[2 * !il].dims[Tids[1][2 * !il]] * d->site_ss[1].dims[Tids[1][1]];
                        const int M1 = d->site_ss[2 * !il].dims[Tids[0][2 * !il]] * d->site_ss[1].dims[Tids[0][1]];
                        int irr[3][3][T3NS_MAX_SYMMETRIES];
                        for (int j = 0; j < 3; ++j) {
                                memcpy(irr[0][j], irrep_of(&oss0[j], ids[0][j]), sizeof(irr[0][j]));
                                memcpy(irr[1][j], irrep_of(&oss0[j], ids[1][j]), sizeof(irr[1][j]));
                                memcpy(irr[2][j], irrep_of(&oss2[j], ids[2][j]), sizeof(irr[2][j]));
                        }
                        const double pref = oracle_prefactor_adjoint(&irr[0][0][0], il ? '3' : '1', d->sgs, ns) *
                                            oracle_prefactor_pUpdate(&irr[0][0][0], il, d->sgs, ns);
                        /* dgemm order of get_upd_aide */
                        const int order = (int64_t) N1 * M1 * (M0 + N1) > (int64_t) M0 * N1 * (M1 + N0);
                        double* work = malloc(sizeof(double) * ((order ? (size_t) M0 * N1 : (size_t) N0 * M1) + 1));
                        for (int k = 0; k < po->nrops; ++k) {
                                if (po->hss_of_ops[k] != newhss) continue;
                                const int64_t* obb = po->beginblock[k];
                                const int64_t* nbb = d->hss_beginblock[newhss];
                                if (obb[osb2 + 1] == obb[osb2] || nbb[usb2 + 1] == nbb[usb2]) continue;
                                const double* O = po->tel[k] + obb[osb2];
                                double* U = ntel[k] + nbb[usb2];
                                if (order) {
                                        /* work = O T (M0 x N1); U += pref Th^T work */
                                        for (int j = 0; j < N1; ++j)
                                                for (int i = 0; i < M0; ++i) {
                                                        double s = 0;
                                                        for (int m = 0; m < M1; ++m)
                                                                s += O[i + (int64_t) m * M0] * site_el(T, il, M1, N1, m, j);
                                                        work[i + (int64_t) j * M0] = s;
                                                }
                                        for (int j = 0; j < N1; ++j)
                                                for (int i = 0; i < N0; ++i) {
                                                        double s = 0;
                                                        for (int m = 0; m < M0; ++m)
                                                                s += adj_el(Th, il, N0, M0, i, m) * work[m + (int64_t) j * M0];
                                                        U[i + (int64_t) j * N0] += pref * s;
                                                }
                                } else {
                                        /* work = Th^T O (N0 x M1); U += pref work T */
                                        for (int j = 0; j < M1; ++j)
                                                for (int i = 0; i < N0; ++i) {
                                                        double s = 0;
                                                        for (int m = 0; m < M0; ++m)
                                                                s += adj_el(Th, il, N0, M0, i, m) * O[m + (int64_t) j * M0];
                                                        work[i + (int64_t) j * N0] = s;
                                                }
                                        for (int j = 0; j < N1; ++j)
                                                for (int i = 0; i < N0; ++i) {
                                                        double s = 0;
                                                        for (int m = 0; m < M1; ++m)
                                                                s += work[i + (int64_t) m * N0] * site_el(T, il, M1, N1, m, j);
                                                        U[i + (int64_t) j * N0] += pref * s;
                                                }
                                }
                        }
                        free(work);
                }
        }
        /* kick_zero_blocks and output */
        int64_t total = 0, nbbo = 0;
        for (int k = 0; k < po->nrops; ++k) {
                const int h = po->hss_of_ops[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                const int64_t* bb = d->hss_beginblock[h];
                int64_t pos = 0;
                if (out_bb) out_bb[nbbo] = 0;
                for (int b = 0; b < nb; ++b) {
                        int nonzero = 0;
                        for (int64_t j = bb[b]; j < bb[b + 1]; ++j)
                                if (!(fabs(ntel[k][j]) < 1e-20)) { nonzero = 1; break; }
                        const int64_t n = nonzero ? bb[b + 1] - bb[b] : 0;
                        if (out_tel && n) memcpy(out_tel + total + pos, ntel[k] + bb[b], sizeof(double) * n);
                        pos += n;
                        if (out_bb) out_bb[nbbo + b + 1] = pos;
                }
                nbbo += nb + 1;
                total += pos;
        }
        for (int k = 0; k < po->nrops; ++k) free(ntel[k]);
        free(ntel);
        free(usb_to_osb);
        return total;
}

/* ------------------------------------------------------------------------
 * update_rOperators_branching (src/rOperators_bUpdate.c:229-248):
 * prepare_update_branching (292-313), init_uniqueOperators (255-290),
 * make_qntens (315-383), init_instrhelper (423-493),
 * initialize_indexhelper (496-532), update_unique_ops_T3NS (637-677),
 * update_newblock_w_MPO_set (746-785), next_sb_tens (680-718),
 * next_sb_sec_op (720-744), find_block_adj (787-806), ways_to_update /
 * calc_contract / fillin_cinfo / how_to_update (813-1012),
 * update_selected_blocks (1043-1072), get_tels_operators (1074-1101),
 * then sum_unique_rOperators (src/rOperators_init.c:228-265).
 * ---------------------------------------------------------------------- */

enum { B_OPS1, B_OPS2, B_NEWOPS, B_TENS, B_ADJ, B_WORKBRA, B_WORKKET };
enum { B_BRA, B_KET, B_MPO };

static const int b_ways[6][3][3] = {
        {{B_OPS1, B_TENS, B_WORKKET}, {B_OPS2, B_WORKKET, B_WORKBRA}, {B_ADJ, B_WORKBRA, B_NEWOPS}},
        {{B_OPS2, B_TENS, B_WORKKET}, {B_OPS1, B_WORKKET, B_WORKBRA}, {B_ADJ, B_WORKBRA, B_NEWOPS}},
        {{B_OPS1, B_ADJ, B_WORKBRA}, {B_OPS2, B_WORKBRA, B_WORKKET}, {B_WORKKET, B_TENS, B_NEWOPS}},
        {{B_OPS2, B_ADJ, B_WORKBRA}, {B_OPS1, B_WORKBRA, B_WORKKET}, {B_WORKKET, B_TENS, B_NEWOPS}},
        {{B_OPS1, B_TENS, B_WORKKET}, {B_OPS2, B_ADJ, B_WORKBRA}, {B_WORKBRA, B_WORKKET, B_NEWOPS}},
        {{B_OPS2, B_TENS, B_WORKKET}, {B_OPS1, B_ADJ, B_WORKBRA}, {B_WORKBRA, B_WORKKET, B_NEWOPS}}};

struct b_cinfo {
        int tens[3];
        int ta, tb;
        int M, N, K, L;
        int lda, ldb, ldc;
        int64_t stride[3];
};

static void b_get_dims(int t, int* rd, int (*td)[2], int (*wd)[2], const int* id_ops)
{
        if (t < B_TENS) {
                rd[B_BRA] = td[id_ops[t]][B_BRA];
                rd[B_KET] = td[id_ops[t]][B_KET];
        } else if (t < B_WORKBRA) {
                const int bt = t == B_TENS ? B_KET : B_BRA;
                for (int i = 0; i < 3; ++i) rd[i] = td[i][bt];
        } else {
                const int bt = t == B_WORKKET ? B_KET : B_BRA;
                for (int i = 0; i < 3; ++i) rd[i] = wd[i][bt];
        }
}

static const int b_bondz[3][2] = {{1, 2}, {0, 2}, {0, 1}};

static int b_ops_needed(int t, int bc, int btype, const int* d1, const int* d2, int (*wd)[2])
{
        const int* bs = b_bondz[bc];
        if (t == B_WORKBRA || t == B_WORKKET) {
                const int bt = t == B_WORKKET ? B_KET : B_BRA;
                const int r = d1[0] * d1[1] * d2[bs[0]] * d2[bs[1]];
                wd[bs[0]][bt] = d2[bs[0]];
                wd[bs[1]][bt] = d2[bs[1]];
                wd[bc][bt] = d1[!btype];
                return r;
        }
        return d1[bs[0]] * d1[bs[1]] * d1[bc] * d2[bc];
}

static int b_calc(const int (*tens)[3], int (*td)[2], int (*wd)[2], const int* id_ops)
{
        int r = 0;
        for (int i = 0; i < 3; ++i) {
                const int bc = i == 2 ? id_ops[tens[i][2]] : id_ops[tens[i][0]];
                const int btype = tens[i][1] == B_TENS || tens[i][1] == B_WORKKET ? B_KET : B_BRA;
                int d1[3], d2[3];
                b_get_dims(tens[i][0], d1, td, wd, id_ops);
                b_get_dims(tens[i][1], d2, td, wd, id_ops);
                r += b_ops_needed(tens[i][2], bc, btype, d1, d2, wd);
        }
        return r;
}

static void b_fill(struct b_cinfo* ci, const int (*tens)[3], int (*td)[2], int (*wd)[2], const int* id_ops)
{
        for (int i = 0; i < 3; ++i) {
                struct b_cinfo* c = &ci[i];
                memset(c, 0, sizeof(*c));
                const int last = i == 2;
                const int bc = last ? id_ops[tens[i][2]] : id_ops[tens[i][0]];
                const int btype = tens[i][1] == B_TENS || tens[i][1] == B_WORKKET ? B_KET : B_BRA;
                const int* bs = b_bondz[bc];
                int d1[3], d2[3];
                b_get_dims(tens[i][0], d1, td, wd, id_ops);
                b_get_dims(tens[i][1], d2, td, wd, id_ops);
                if (last) {
                        c->ta = bc != 0;
                        c->tb = bc == 0;
                        c->M = d1[bc];
                        c->N = d2[bc];
                        if (bc == 1) {
                                c->K = d2[bs[0]];
                                c->L = d2[bs[1]];
                                c->stride[0] = (int64_t) c->M * c->K;
                                c->stride[1] = (int64_t) c->N * c->K;
                        } else {
                                c->K = d1[bs[0]] * d1[bs[1]];
                                c->L = 1;
                        }
                        c->tens[0] = tens[i][0];
                        c->tens[1] = tens[i][1];
                        c->tens[2] = tens[i][2];
                } else if (bc == 0) {
                        c->ta = btype != B_KET;
                        c->tb = 0;
                        c->M = d1[!btype];
                        c->N = d2[bs[0]] * d2[bs[1]];
                        c->K = d1[btype];
                        c->L = 1;
                        c->tens[0] = tens[i][0];
                        c->tens[1] = tens[i][1];
                        c->tens[2] = tens[i][2];
                } else if (bc == 1) {
                        c->ta = 0;
                        c->tb = btype == B_KET;
                        c->M = d2[bs[0]];
                        c->N = d1[!btype];
                        c->K = d1[btype];
                        c->L = d2[bs[1]];
                        c->stride[0] = (int64_t) c->K * c->M;
                        c->stride[2] = (int64_t) c->N * c->M;
                        c->tens[0] = tens[i][1];
                        c->tens[1] = tens[i][0];
                        c->tens[2] = tens[i][2];
                } else {
                        c->ta = 0;
                        c->tb = btype != B_BRA;
                        c->M = d2[bs[0]] * d2[bs[1]];
                        c->N = d1[!btype];
                        c->K = d1[btype];
                        c->L = 1;
                        c->tens[0] = tens[i][1];
                        c->tens[1] = tens[i][0];
                        c->tens[2] = tens[i][2];
                }
                c->lda = !c->ta ? c->M : c->K;
                c->ldb = !c->tb ? c->K : c->N;
                c->ldc = c->M;
        }
}

/* -1: the cheapest of the six contraction orders, as the reference;
 * 0..5: that order for every block (for the order-independence check) */
static int b_forced_way = -1;

void oracle_set_bupdate_order(int way)
{
        b_forced_way = way;
}

static int64_t b_block_size(const t3ns_rops_desc* r, int op, int b)
{
        return r->beginblock[op][b + 1] - r->beginblock[op][b];
}

int64_t oracle_update_branching(const t3ns_bupdate_desc* d, const t3ns_rops_desc* ops,
                                int64_t* out_bb, double* out_tel)
{
        const int uc = d->update_case;
        const int H = d->mpo.nrSecs;
        const int id_ops[3] = {uc == 0 ? 1 : 0, uc == 2 ? 1 : 2, uc};
        const int looptype = uc == 0;
        int maxdims[3][3];
        for (int i = 0; i < 3; ++i) {
                maxdims[i][B_BRA] = maxdims[i][B_KET] = d->site_ss[i].nrSecs;
                maxdims[i][B_MPO] = H;
        }
        /* unique operators */
        int nu = 0;
        int* run_first = malloc(sizeof(int) * (d->nr_instr + 1));
        int* run_of = malloc(sizeof(int) * (d->nr_instr + 1));
        for (int i = 0; i < d->nr_instr; ++i) {
                const int* a = d->instr + 3 * i;
                const int* b = d->instr + 3 * (i - 1);
                if (i == 0 || a[0] != b[0] || a[1] != b[1] || d->hss_of_new[a[2]] != d->hss_of_new[b[2]])
                        run_first[nu++] = i;
                run_of[i] = nu - 1;
        }
        double** utel = malloc(sizeof(double*) * (nu + 1));
        int* u_hss = malloc(sizeof(int) * (nu + 1));
        for (int u = 0; u < nu; ++u) {
                const int h = d->hss_of_new[d->instr[3 * run_first[u] + 2]];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                u_hss[u] = h;
                utel[u] = calloc((size_t) d->hss_beginblock[h][nb] + 1, sizeof(double));
        }
        /* make_qntens: tensor blocks grouped by the two kets the loop fixes */
        const int64_t divide = looptype ? (int64_t) maxdims[0][B_KET] * maxdims[1][B_KET] : maxdims[0][B_KET];
        /* (the groups are searched by linear scan below, in block order) */

        const int first = looptype ? B_OPS1 : B_OPS2, second = looptype ? B_OPS2 : B_OPS1;
        const int N = d->begin_blocks_of_hss[d->nrhss];
        for (int new_sb = 0; new_sb < N; ++new_sb) {
                int id[3][3];
                int64_t qn = d->qnumbers[new_sb];
                const int nb_ = id_ops[B_NEWOPS];
                id[nb_][B_BRA] = (int) (qn % maxdims[nb_][B_BRA]);
                qn /= maxdims[nb_][B_BRA];
                id[nb_][B_KET] = (int) (qn % maxdims[nb_][B_KET]);
                id[nb_][B_MPO] = (int) (qn / maxdims[nb_][B_KET]);
                const int hn = id[nb_][B_MPO];
                const int sb_new = new_sb - d->begin_blocks_of_hss[hn];
                for (int p = d->tprod_begin[hn]; p < d->tprod_begin[hn + 1]; ++p) {
                        const int hss_ops[2] = {d->tprods[2 * p], d->tprods[2 * p + 1]};
                        const t3ns_rops_desc* F = &ops[first];
                        const t3ns_rops_desc* S = &ops[second];
                        const int bf = id_ops[first], bs = id_ops[second];
                        const int f0 = F->begin_blocks_of_hss[hss_ops[first]];
                        const int nbf = F->begin_blocks_of_hss[hss_ops[first] + 1] - f0;
                        for (int jf = 0; jf < nbf; ++jf) {
                                int64_t q = F->qnumbers[f0 + jf];
                                id[bf][B_BRA] = (int) (q % maxdims[bf][B_BRA]);
                                q /= maxdims[bf][B_BRA];
                                id[bf][B_KET] = (int) (q % maxdims[bf][B_KET]);
                                id[bf][B_MPO] = (int) (q / maxdims[bf][B_KET]);
                                id[bs][B_MPO] = hss_ops[second];
                                const int64_t match = looptype ? id[0][B_KET] + (int64_t) id[1][B_KET] * maxdims[0][B_KET]
                                                               : id[1][B_KET] + (int64_t) id[2][B_KET] * maxdims[1][B_KET];
                                for (int tb = 0; tb < d->T_nrblocks; ++tb) {
                                        const int64_t tq = d->T_qnumbers[tb];
                                        if ((looptype ? tq % divide : tq / divide) != match) continue;
                                        id[bs][B_KET] = (int) (looptype ? tq / divide : tq % divide);
                                        const int s0 = S->begin_blocks_of_hss[hss_ops[second]];
                                        const int nbs = S->begin_blocks_of_hss[hss_ops[second] + 1] - s0;
                                        for (int js = 0; js < nbs; ++js) {
                                                const int64_t sq = S->qnumbers[s0 + js];
                                                if (sq / maxdims[bs][B_BRA] != id[bs][B_KET] + (int64_t) hss_ops[second] * maxdims[bs][B_KET])
                                                        continue;
                                                id[bs][B_BRA] = (int) (sq % maxdims[bs][B_BRA]);
                                                /* find_block_adj */
                                                int64_t aq = 0;
                                                for (int i = 2; i >= 0; --i) aq = aq * maxdims[i][B_BRA] + id[i][B_BRA];
                                                const int ab = bsearch_qn(aq, d->T_qnumbers, d->T_nrblocks);
                                                if (ab < 0) continue;
                                                int sb_op[3];
                                                sb_op[first] = jf;
                                                sb_op[second] = js;
                                                sb_op[B_NEWOPS] = sb_new;
                                                /* update_selected_blocks */
                                                int irr[3][3][T3NS_MAX_SYMMETRIES];
                                                int td[3][2];
                                                for (int i = 0; i < 3; ++i) {
                                                        memcpy(irr[i][0], irrep_of(&d->site_ss[i], id[i][B_BRA]), sizeof(irr[i][0]));
                                                        memcpy(irr[i][1], irrep_of(&d->site_ss[i], id[i][B_KET]), sizeof(irr[i][1]));
                                                        memcpy(irr[i][2], irrep_of(&d->mpo, id[i][B_MPO]), sizeof(irr[i][2]));
                                                        td[i][B_BRA] = d->site_ss[i].dims[id[i][B_BRA]];
                                                        td[i][B_KET] = d->site_ss[i].dims[id[i][B_KET]];
                                                }
                                                const double pref = oracle_prefactor_bUpdate(&irr[0][0][0], uc, d->sgs, d->nrSyms);
                                                struct b_cinfo ci[3];
                                                int wsz[2] = {0, 0};
                                                int best = 0;
                                                for (int w = 0; w < 6; ++w) {
                                                        int wd[3][2];
                                                        const int cost = b_calc(b_ways[w], td, wd, id_ops);
                                                        if (b_forced_way >= 0 ? w == b_forced_way
                                                                              : (w == 0 || best > cost)) {
                                                                best = cost;
                                                                b_fill(ci, b_ways[w], td, wd, id_ops);
                                                                wsz[0] = wd[0][B_BRA] * wd[1][B_BRA] * wd[2][B_BRA];
                                                                wsz[1] = wd[0][B_KET] * wd[1][B_KET] * wd[2][B_KET];
                                                        }
                                                }
                                                double* tel[7] = {0};
                                                tel[B_TENS] = (double*) d->T_tel + d->T_beginblock[tb];
                                                tel[B_ADJ] = (double*) d->T_tel + d->T_beginblock[ab];
                                                tel[B_WORKBRA] = malloc(sizeof(double) * ((size_t) wsz[0] + 1));
                                                tel[B_WORKKET] = malloc(sizeof(double) * ((size_t) wsz[1] + 1));
                                                /* find_matching_instr: unique runs of this MPO combination */
                                                for (int u = 0; u < nu; ++u) {
                                                        const int* ins = d->instr + 3 * run_first[u];
                                                        if (ops[0].hss_of_ops[ins[0]] != hss_ops[0] ||
                                                            ops[1].hss_of_ops[ins[1]] != hss_ops[1] ||
                                                            d->hss_of_new[ins[2]] != hn)
                                                                continue;
                                                        if (b_block_size(&ops[0], ins[0], sb_op[B_OPS1]) == 0 ||
                                                            b_block_size(&ops[1], ins[1], sb_op[B_OPS2]) == 0)
                                                                continue;
                                                        const int64_t* ubb = d->hss_beginblock[hn];
                                                        if (ubb[sb_new + 1] == ubb[sb_new]) continue;
                                                        tel[B_OPS1] = (double*) ops[0].tel[ins[0]] + ops[0].beginblock[ins[0]][sb_op[B_OPS1]];
                                                        tel[B_OPS2] = (double*) ops[1].tel[ins[1]] + ops[1].beginblock[ins[1]][sb_op[B_OPS2]];
                                                        tel[B_NEWOPS] = utel[u] + ubb[sb_new];
                                                        for (int c = 0; c < 3; ++c) {
                                                                const struct b_cinfo* x = &ci[c];
                                                                for (int l = 0; l < x->L; ++l)
                                                                        oracle_dgemm_batched(x->ta, x->tb, x->M, x->N, x->K, 1,
                                                                                             c == 2 ? pref : 1.0,
                                                                                             tel[x->tens[0]] + l * x->stride[0], x->lda, 0,
                                                                                             tel[x->tens[1]] + l * x->stride[1], x->ldb, 0,
                                                                                             c == 2 ? 1.0 : 0.0,
                                                                                             tel[x->tens[2]] + l * x->stride[2], x->ldc, 0);
                                                        }
                                                }
                                                free(tel[B_WORKBRA]);
                                                free(tel[B_WORKKET]);
                                        }
                                }
                        }
                }
        }
        /* sum_unique_rOperators */
        int nops = 0;
        for (int i = 0; i < d->nr_instr; ++i)
                if (d->instr[3 * i + 2] + 1 > nops) nops = d->instr[3 * i + 2] + 1;
        double** ntel = malloc(sizeof(double*) * (nops + 1));
        for (int k = 0; k < nops; ++k) {
                const int h = d->hss_of_new[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                ntel[k] = calloc((size_t) d->hss_beginblock[h][nb] + 1, sizeof(double));
        }
        for (int i = 0; i < d->nr_instr; ++i) {
                const int k = d->instr[3 * i + 2];
                const int h = d->hss_of_new[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                const int64_t n = d->hss_beginblock[h][nb];
                for (int64_t j = 0; j < n; ++j) ntel[k][j] += d->pref[i] * utel[run_of[i]][j];
        }
        int64_t total = 0, nbbo = 0;
        for (int k = 0; k < nops; ++k) {
                const int h = d->hss_of_new[k];
                const int nb = d->begin_blocks_of_hss[h + 1] - d->begin_blocks_of_hss[h];
                const int64_t* bb = d->hss_beginblock[h];
                int64_t pos = 0;
                if (out_bb) out_bb[nbbo] = 0;
                for (int b = 0; b < nb; ++b) {
                        int nonzero = 0;
                        for (int64_t j = bb[b]; j < bb[b + 1]; ++j)
                                if (!(fabs(ntel[k][j]) < 1e-20)) { nonzero = 1; break; }
                        const int64_t n = nonzero ? bb[b + 1] - bb[b] : 0;
                        if (out_tel && n) memcpy(out_tel + total + pos, ntel[k] + bb[b], sizeof(double) * n);
                        pos += n;
                        if (out_bb) out_bb[nbbo + b + 1] = pos;
                }
                nbbo += nb + 1;
                total += pos;
        }
        for (int u = 0; u < nu; ++u) free(utel[u]);
        for (int k = 0; k < nops; ++k) free(ntel[k]);
        free(utel);
        free(ntel);
        free(u_hss);
        free(run_first);
        free(run_of);
        return total;
}
