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
