/* ORACLE -- test infrastructure, not product code.
 *
 * Plain-C restatement of makesiteTensor's contraction of one-site tensors
 * into the multisite tensor, following the reference step by step:
 *   the inner site and its connections (init_md, src/siteTensor_init.c:
 *       192-260),
 *   init_makeinfo (609-649): per multisite block, the one-site blocks
 *       through the translated quantum numbers,
 *   init_contractinfo (651-760): the GEMM sequence, including its
 *       transposed and batched forms,
 *   do_contract (src/sparseblocks.c:133-158; a batch count of 0 still
 *       executes one product).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int sector_index(const int* irrep, const t3ns_symsecs_desc* ss)
{
        for (int i = 0; i < ss->nrSecs; ++i)
                if (!memcmp(ss->irreps + (size_t) i * T3NS_MAX_SYMMETRIES, irrep,
                            sizeof(int) * T3NS_MAX_SYMMETRIES))
                        return i;
        return -1;
}

static int find_qn(int64_t v, const int64_t* a, int n)
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

enum { SITE1, SITE2, SITE3, SITE4, RESULT, WORK1, WORK2 };

struct cinfo {
        int tens[3];
        int ta, tb;
        int M, N, K, L;
        int lda, ldb, ldc;
        int64_t stride[3];
};

int oracle_make_site_tensor(const t3ns_makesite_desc* d, double* tel)
{
        const int ns = d->nrsites, in = d->innersite;
        /* connected[leg of the inner site] = (site, its leg) */
        int connected[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
        int nint = 0;
        for (int j = 0; j < 3; ++j)
                for (int s = 0; s < ns; ++s) {
                        if (s == in) continue;
                        for (int k = 0; k < 3; ++k)
                                if (d->new_ss[s][k].bond == d->new_ss[in][j].bond) {
                                        connected[j][0] = s;
                                        connected[j][1] = k;
                                        ++nint;
                                }
                }
        if (nint != ns - 1) return -1;
        const int64_t total = d->beginblock[d->nrblocks];
        memset(tel, 0, sizeof(double) * (size_t) total);
        for (int sb = 0; sb < d->nrblocks; ++sb) {
                /* init_makeinfo */
                double* t[7] = {0};
                int odim[4][3];
                int valid = 1;
                for (int i = 0; i < ns && valid; ++i) {
                        const t3ns_symsecs_desc* nss = d->new_ss[i];
                        const t3ns_symsecs_desc* oss = d->old_ss[i];
                        const int64_t qn = d->qnumbers[(size_t) sb * ns + i];
                        const int64_t n0 = nss[0].nrSecs, n1 = nss[1].nrSecs;
                        const int nid[3] = {(int) (qn % n0), (int) ((qn / n0) % n1), (int) (qn / (n0 * n1))};
                        int oid[3];
                        for (int k = 0; k < 3; ++k) {
                                oid[k] = sector_index(nss[k].irreps + (size_t) nid[k] * T3NS_MAX_SYMMETRIES, &oss[k]);
                                if (oid[k] < 0) valid = 0;
                        }
                        if (!valid) break;
                        const int64_t oqn = oid[0] + (int64_t) oid[1] * oss[0].nrSecs +
                                            (int64_t) oid[2] * oss[0].nrSecs * oss[1].nrSecs;
                        const int csb = find_qn(oqn, d->site_qnumbers[i], d->site_nrblocks[i]);
                        if (csb < 0 || d->site_beginblock[i][csb + 1] == d->site_beginblock[i][csb]) {
                                valid = 0;
                                break;
                        }
                        t[SITE1 + i] = (double*) d->site_tel[i] + d->site_beginblock[i][csb];
                        for (int k = 0; k < 3; ++k) odim[i][k] = oss[k].dims[oid[k]];
                }
                if (!valid) continue;
                t[RESULT] = tel + d->beginblock[sb];

                /* init_contractinfo with the internal bonds in inner-leg order */
                const int result[4] = {SITE1 + in, WORK1, WORK2, RESULT};
                int innerdim[3] = {odim[in][0], odim[in][1], odim[in][2]};
                int one_added = 0, i = 0;
                struct cinfo ci[3];
                for (int bi = 0; bi < 3; ++bi) {
                        if (connected[bi][0] < 0) continue;
                        struct cinfo* c = &ci[i];
                        memset(c, 0, sizeof(*c));
                        const int site = connected[bi][0], bond = connected[bi][1];
                        if (site > in) {
                                c->tens[0] = result[i];
                                c->ta = 0;
                                c->M = innerdim[0] * innerdim[1];
                                c->K = innerdim[2];
                                c->lda = c->M;
                                c->tens[1] = SITE1 + site;
                                c->tb = 0;
                                c->N = odim[site][1] * odim[site][2];
                                c->ldb = c->K;
                                c->L = 0;
                                innerdim[bi] = c->N;
                        } else {
                                c->tens[0] = SITE1 + site;
                                c->ta = bond == 0;
                                c->M = odim[site][1] * odim[site][2 * (bond == 0)];
                                c->K = odim[site][bond];
                                c->lda = bond == 0 ? c->K : c->M;
                                c->stride[0] = 0;
                                c->tens[1] = result[i];
                                c->tb = bi != 0;
                                if (bi == 0) {
                                        c->N = innerdim[1] * innerdim[2];
                                        c->L = 1;
                                        c->ldb = c->K;
                                        one_added = 1;
                                } else if (bi == 1) {
                                        c->N = innerdim[0];
                                        c->L = innerdim[2];
                                        c->ldb = c->N;
                                        c->stride[1] = (int64_t) c->N * c->K;
                                } else {
                                        c->N = innerdim[0] * innerdim[1];
                                        c->L = 1;
                                        c->ldb = c->N;
                                }
                                if (one_added && bi == 1) {
                                        c->tens[0] = result[i];
                                        c->ta = 0;
                                        c->M = innerdim[0];
                                        c->K = innerdim[1];
                                        c->L = innerdim[2];
                                        c->lda = c->M;
                                        c->tens[1] = SITE1 + site;
                                        c->tb = 1;
                                        c->N = odim[site][0] * odim[site][1];
                                        c->ldb = c->N;
                                        c->stride[0] = (int64_t) c->M * c->K;
                                        c->stride[1] = 0;
                                        innerdim[bi] = c->N;
                                } else {
                                        innerdim[bi] = c->M;
                                }
                        }
                        c->tens[2] = result[i + 1];
                        c->ldc = c->M;
                        c->stride[2] = (int64_t) c->N * c->M;
                        ++i;
                }
                const int nc = i;
                for (int k = 0; k < nc; ++k) {
                        if (k == nc - 1) {
                                ci[k].tens[2] = RESULT;
                        } else {
                                const int L = ci[k].L > 0 ? ci[k].L : 1;
                                t[ci[k].tens[2]] = malloc(sizeof(double) * ((size_t) ci[k].M * ci[k].N * L + 1));
                        }
                }
                for (int k = 0; k < nc; ++k) {
                        const struct cinfo* c = &ci[k];
                        const int L = c->L > 0 ? c->L : 1;
                        for (int l = 0; l < L; ++l)
                                oracle_dgemm_batched(c->ta, c->tb, c->M, c->N, c->K, 1, 1.0,
                                                     t[c->tens[0]] + l * c->stride[0], c->lda, 0,
                                                     t[c->tens[1]] + l * c->stride[1], c->ldb, 0, 0.0,
                                                     t[c->tens[2]] + l * c->stride[2], c->ldc, 0);
                }
                free(t[WORK1]);
                free(t[WORK2]);
        }
        return 0;
}
