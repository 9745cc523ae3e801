/* ORACLE -- test infrastructure, not product code.
 *
 * Plain-C restatement of decompose_siteTensor's two paths.  First
 * split_of_site (src/siteTensor_decompose.c:
 * 1403-1444): init_splitted_tens (633-660), make_r_count_svdinfos and
 * make_svdinfos (680-842), SVD_copy_to_mem (955-987), svdblocks (1050-1072;
 * LAPACK dgesdd replaced by a sequential one-sided Jacobi SVD written
 * here), selectS (1196-1289), init_UV_tensors_and_change_symsec
 * (1290-1330), SVD_copy_from_mem (989-1048), adapt_UV_tensors_and_kick_
 * empties (1332-1401), norm_tensor (src/siteTensor_misc.c:244-250).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NS T3NS_MAX_SYMMETRIES

static void idx3(int* id, int64_t qn, const t3ns_symsecs_desc* ss)
{
        const int64_t n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
        id[0] = (int) (qn % n0);
        id[1] = (int) ((qn / n0) % n1);
        id[2] = (int) (qn / (n0 * n1));
}

static int64_t size3(const t3ns_symsecs_desc* ss, int64_t qn, int from, int to)
{
        int id[3];
        idx3(id, qn, ss);
        int64_t r = 1;
        for (int k = from; k < to; ++k) r *= ss[k].dims[id[k]];
        return r;
}

static int g_tuple_n;
static int cmp_tuple(const void* a, const void* b)
{
        const int64_t* x = a;
        const int64_t* y = b;
        for (int i = g_tuple_n - 1; i >= 0; --i)
                if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
        return 0;
}

/* sorted unique tuples of the chosen sites of A (init_splitted_tens) */
static int64_t* split_tuples(const t3ns_split_desc* d, const int* incl, int n, int* count)
{
        int64_t* t = malloc(sizeof(int64_t) * ((size_t) d->nrblocks * n + 1));
        for (int b = 0; b < d->nrblocks; ++b)
                for (int j = 0; j < n; ++j) t[(size_t) b * n + j] = d->qnumbers[(size_t) b * d->nrsites + incl[j]];
        g_tuple_n = n;
        qsort(t, d->nrblocks, sizeof(int64_t) * n, cmp_tuple);
        int cnt = d->nrblocks ? 1 : 0;
        for (int b = 1; b < d->nrblocks; ++b)
                if (cmp_tuple(t + (size_t) b * n, t + (size_t) (cnt - 1) * n) != 0) {
                        memmove(t + (size_t) cnt * n, t + (size_t) b * n, sizeof(int64_t) * n);
                        ++cnt;
                }
        *count = cnt;
        return t;
}

/* One-sided Jacobi SVD of an m x n column-major matrix (m >= n), cyclic
 * pair order.  B becomes U * Sigma (unsorted), V (n x n) the right vectors. */
static void jacobi(double* B, int m, int n, double* V)
{
        for (int i = 0; i < n * n; ++i) V[i] = (i % (n + 1) == 0);
        const double tol = 1e-15 * sqrt((double) m);
        for (int sweep = 0; sweep < 60; ++sweep) {
                int rot = 0;
                for (int p = 0; p < n; ++p)
                        for (int q = p + 1; q < n; ++q) {
                                double a = 0, b = 0, g = 0;
                                double* x = B + (size_t) p * m;
                                double* y = B + (size_t) q * m;
                                for (int i = 0; i < m; ++i) {
                                        a += x[i] * x[i];
                                        b += y[i] * y[i];
                                        g += x[i] * y[i];
                                }
                                if (g == 0 || fabs(g) <= tol * sqrt(a * b)) continue;
                                rot = 1;
                                const double zeta = (b - a) / (2 * g);
                                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                                const double c = 1 / sqrt(1 + t * t), s = c * t;
                                for (int i = 0; i < m; ++i) {
                                        const double u = x[i], v = y[i];
                                        x[i] = c * u - s * v;
                                        y[i] = s * u + c * v;
                                }
                                double* vp = V + (size_t) p * n;
                                double* vq = V + (size_t) q * n;
                                for (int i = 0; i < n; ++i) {
                                        const double u = vp[i], v = vq[i];
                                        vp[i] = c * u - s * v;
                                        vq[i] = s * u + c * v;
                                }
                        }
                if (!rot) break;
        }
}

/* Rows t of VT (k x n, row-major over the stride k) whose singular value is
 * at most 1e-12 of the largest carry no direction (the rotated column they
 * come from is rounding noise, or zero); dgesdd still returns orthonormal
 * rows there.  They are replaced, in order, by the first unit vector e_c
 * (c = 0, 1, ...) that keeps at least half its norm after two
 * Gram-Schmidt passes against every other good or completed row. */
static void complete_rows(double* VT, const double* S, int k, int n)
{
        double smax = 0;
        for (int t = 0; t < k; ++t) smax = S[t] > smax ? S[t] : smax;
        const double thr = 1e-12 * smax;
        char* ok = malloc(k + 1);
        for (int t = 0; t < k; ++t) ok[t] = S[t] > thr;
        double* v = malloc(sizeof(double) * (n + 1));
        for (int t = 0; t < k; ++t) {
                if (ok[t]) continue;
                for (int c = 0; c < n; ++c) {
                        for (int i = 0; i < n; ++i) v[i] = i == c;
                        for (int pass = 0; pass < 2; ++pass)
                                for (int r = 0; r < k; ++r) {
                                        if (!ok[r]) continue;
                                        double d = 0;
                                        for (int i = 0; i < n; ++i) d += VT[r + (size_t) i * k] * v[i];
                                        for (int i = 0; i < n; ++i) v[i] -= d * VT[r + (size_t) i * k];
                                }
                        double nr = 0;
                        for (int i = 0; i < n; ++i) nr += v[i] * v[i];
                        nr = sqrt(nr);
                        if (nr > 0.5) {
                                for (int i = 0; i < n; ++i) VT[t + (size_t) i * k] = v[i] / nr;
                                ok[t] = 1;
                                break;
                        }
                }
        }
        free(v);
        free(ok);
}

/* dgesdd 'S' semantics: US = U * diag(S) (m x k), VT (k x n), S descending, k = min(m, n) */
static void svd(const double* A, int m, int n, double* S, double* US, double* VT)
{
        const int tr = m < n;
        const int mm = tr ? n : m, nn = tr ? m : n;
        double* B = malloc(sizeof(double) * ((size_t) mm * nn + 1));
        double* V = malloc(sizeof(double) * ((size_t) nn * nn + 1));
        for (int i = 0; i < m; ++i)
                for (int j = 0; j < n; ++j)
                        if (tr) B[j + (size_t) i * n] = A[i + (size_t) j * m];
                        else B[i + (size_t) j * m] = A[i + (size_t) j * m];
        jacobi(B, mm, nn, V);
        double* sv = malloc(sizeof(double) * (nn + 1));
        int* ord = malloc(sizeof(int) * (nn + 1));
        for (int j = 0; j < nn; ++j) {
                double a = 0;
                for (int i = 0; i < mm; ++i) a += B[i + (size_t) j * mm] * B[i + (size_t) j * mm];
                sv[j] = sqrt(a);
                ord[j] = j;
        }
        for (int i = 1; i < nn; ++i)  /* insertion sort, descending, stable */
                for (int j = i; j > 0 && sv[ord[j]] > sv[ord[j - 1]]; --j) {
                        const int t = ord[j];
                        ord[j] = ord[j - 1];
                        ord[j - 1] = t;
                }
        for (int t = 0; t < nn; ++t) {
                const int j = ord[t];
                S[t] = sv[j];
                if (!tr) {
                        for (int i = 0; i < m; ++i) US[i + (size_t) t * m] = B[i + (size_t) j * m];
                        for (int i = 0; i < n; ++i) VT[t + (size_t) i * nn] = V[i + (size_t) j * n];
                } else {
                        /* A = V_B Sigma U_B^T */
                        for (int i = 0; i < m; ++i) US[i + (size_t) t * m] = V[i + (size_t) j * m] * sv[j];
                        for (int i = 0; i < n; ++i)
                                VT[t + (size_t) i * nn] = sv[j] > 0 ? B[i + (size_t) j * n] / sv[j] : 0;
                }
        }
        if (tr) complete_rows(VT, S, nn, n);
        free(B);
        free(V);
        free(sv);
        free(ord);
}

static double vn_entropy(double s)
{
        const double w = s * s;
        return w < 1e-20 ? 0 : -w * log(w);
}

static int multiplicity_of(const int* irr, const int* sgs, int nrsy)
{
        int r = 1;
        for (int i = 0; i < nrsy; ++i)
                if (sgs[i] == 2 /* SU2 */) r *= irr[i] + 1;
        return r;
}

int oracle_split_site(const t3ns_split_desc* d, t3ns_split_view* out)
{
        const int ns = d->nrsites, sv = d->siteV;
        /* cut bond */
        /* good_site_to_split (1074-1118): exactly one bond to the rest */
        int id_bond = -1, id_csite = -1, id_cbond = -1, links = 0;
        for (int i = 0; i < ns; ++i) {
                if (i == sv) continue;
                for (int j = 0; j < 3; ++j)
                        for (int k = 0; k < 3; ++k)
                                if (d->ss[sv][k].bond == d->ss[i][j].bond) {
                                        id_csite = i;
                                        id_cbond = j;
                                        id_bond = k;
                                        ++links;
                                }
        }
        if (links != 1) return -1;
        int incl[4], nincl = 0;
        for (int i = 0; i < ns; ++i)
                if (i != sv) incl[nincl++] = i;
        int nU, nV;
        int64_t* Uq = split_tuples(d, incl, nincl, &nU);
        int64_t* Vq = split_tuples(d, &sv, 1, &nV);
        const t3ns_symsecs_desc* sV = d->ss[sv];
        const t3ns_symsecs_desc* cut = &sV[id_bond];
        const int nsec = cut->nrSecs;
        int64_t divV = 1;
        for (int i = 0; i < id_bond; ++i) divV *= sV[i].nrSecs;
        const int csU = id_csite - (id_csite > sv);
        int64_t divC = 1;
        for (int i = 0; i < id_cbond; ++i) divC *= d->ss[id_csite][i].nrSecs;
        const int64_t modC = d->ss[id_csite][id_cbond].nrSecs;

        /* per sector: U rows (Mstart), V columns (Nstart) */
        int64_t* Mrow = calloc((size_t) nU + 1, sizeof(int64_t));   /* row offset of U block */
        int64_t* Ncol = calloc((size_t) nV + 1, sizeof(int64_t));
        int* Usec = malloc(sizeof(int) * (nU + 1));
        int* Vsec = malloc(sizeof(int) * (nV + 1));
        int64_t* M = calloc(nsec + 1, sizeof(int64_t));
        int64_t* N = calloc(nsec + 1, sizeof(int64_t));
        for (int u = 0; u < nU; ++u) {
                const int s = (int) ((Uq[(size_t) u * nincl + csU] / divC) % modC);
                int64_t sz = 1;
                for (int i = 0; i < nincl; ++i) sz *= size3(d->ss[incl[i]], Uq[(size_t) u * nincl + i], 0, 3);
                Usec[u] = s;
                Mrow[u] = M[s];
                M[s] += sz;
        }
        for (int v = 0; v < nV; ++v) {
                const int s = (int) ((Vq[v] / divV) % cut->nrSecs);
                Vsec[v] = s;
                Ncol[v] = N[s];
                N[s] += size3(sV, Vq[v], 0, 3);
        }
        /* sector matrices (SVD_copy_to_mem) */
        double** mat = calloc(nsec + 1, sizeof(double*));
        for (int s = 0; s < nsec; ++s) mat[s] = calloc((size_t) M[s] * N[s] + 1, sizeof(double));
        for (int b = 0; b < d->nrblocks; ++b) {
                const int64_t* qn = d->qnumbers + (size_t) b * ns;
                int64_t tu[4];
                for (int i = 0; i < nincl; ++i) tu[i] = qn[incl[i]];
                g_tuple_n = nincl;
                const int64_t* fu = bsearch(tu, Uq, nU, sizeof(int64_t) * nincl, cmp_tuple);
                g_tuple_n = 1;
                const int64_t* fv = bsearch(&qn[sv], Vq, nV, sizeof(int64_t), cmp_tuple);
                const int u = (int) ((fu - Uq) / nincl), v = (int) (fv - Vq);
                const int s = Vsec[v];
                int64_t dm[3] = {1, 1, 1};
                int c = 0;
                for (int i = 0; i < ns; ++i) {
                        const int64_t sz = size3(d->ss[i], qn[i], 0, 3);
                        if (i == sv) {
                                dm[1] = sz;
                                c = 2;
                        } else {
                                dm[c] *= sz;
                        }
                }
                const double* src = d->tel + d->beginblock[b];
                for (int64_t i2 = 0; i2 < dm[2]; ++i2)
                        for (int64_t i1 = 0; i1 < dm[1]; ++i1)
                                for (int64_t i0 = 0; i0 < dm[0]; ++i0)
                                        mat[s][(Mrow[u] + i0 + i2 * dm[0]) + (Ncol[v] + i1) * M[s]] =
                                                src[i0 + i1 * dm[0] + i2 * dm[0] * dm[1]];
        }
        /* SVDs */
        int* nsing = calloc(nsec + 1, sizeof(int));
        double** S = calloc(nsec + 1, sizeof(double*));
        double** US = calloc(nsec + 1, sizeof(double*));
        double** VT = calloc(nsec + 1, sizeof(double*));
        int total = 0;
        for (int s = 0; s < nsec; ++s) {
                const int k = (int) (M[s] < N[s] ? M[s] : N[s]);
                nsing[s] = k;
                total += k;
                S[s] = malloc(sizeof(double) * (k + 1));
                US[s] = malloc(sizeof(double) * ((size_t) M[s] * k + 1));
                VT[s] = malloc(sizeof(double) * ((size_t) k * N[s] + 1));
                if (k) svd(mat[s], (int) M[s], (int) N[s], S[s], US[s], VT[s]);
        }
        /* selectS */
        int* mult = malloc(sizeof(int) * (nsec + 1));
        double norm0 = 0, ent0 = 0;
        for (int s = 0; s < nsec; ++s) {
                mult[s] = multiplicity_of(cut->irreps + (size_t) s * NS, d->sgs, d->nrSyms);
                const double mf = 1 / sqrt((double) mult[s]);
                double n = 0, e = 0;
                for (int t = 0; t < nsing[s]; ++t) {
                        n += (S[s][t] * mf) * (S[s][t] * mf);
                        e += vn_entropy(S[s][t] * mf);
                }
                norm0 += mult[s] * n;
                ent0 += mult[s] * e;
        }
        double* all = malloc(sizeof(double) * (total + 1));
        int* orig = malloc(sizeof(int) * (total + 1));
        int* order = malloc(sizeof(int) * (total + 1));
        int c = 0;
        for (int s = 0; s < nsec; ++s)
                for (int t = 0; t < nsing[s]; ++t) {
                        all[c] = S[s][t];
                        orig[c] = s;
                        order[c] = c;
                        ++c;
                }
        for (int i = 1; i < total; ++i)  /* ascending, stable */
                for (int j = i; j > 0 && all[order[j]] < all[order[j - 1]]; --j) {
                        const int t = order[j];
                        order[j] = order[j - 1];
                        order[j - 1] = t;
                }
        int* kept = calloc(nsec + 1, sizeof(int));
        const int runupto = d->maxD < total ? d->maxD : total;
        const int minimalto = d->minD < total ? d->minD : total;
        double norm1 = 0, ent1 = 0;
        int ss = 0;
        for (; ss < runupto; ++ss) {
                const int k = order[total - ss - 1];
                const double sg = all[k] / sqrt((double) mult[orig[k]]);
                norm1 += mult[orig[k]] * sg * sg;
                ent1 += mult[orig[k]] * vn_entropy(sg);
                ++kept[orig[k]];
                if (ss < minimalto) continue;
                const int ok = d->truncType == 'N' ? norm0 - norm1 < d->truncerr
                                                   : ent0 - ent1 / norm1 - log(norm1) < d->truncerr;
                if (ok) {
                        ++ss;
                        break;
                }
        }
        ent1 = ent1 / norm1 + log(norm1);

        /* U and V blocks (SVD_copy_from_mem), then the kept sectors only */
        int* newid = malloc(sizeof(int) * (nsec + 1));
        int cnt = 0;
        for (int s = 0; s < nsec; ++s) newid[s] = kept[s] ? cnt++ : -1;
        int64_t* ub = calloc((size_t) nU + 1, sizeof(int64_t));
        int64_t* vb = calloc((size_t) nV + 1, sizeof(int64_t));
        int64_t usz = 0, vsz = 0;
        int nUo = 0, nVo = 0;
        for (int u = 0; u < nU; ++u) {
                int64_t rows = 1;
                for (int i = 0; i < nincl; ++i) rows *= size3(d->ss[incl[i]], Uq[(size_t) u * nincl + i], 0, 3);
                usz += rows * kept[Usec[u]];
                nUo += kept[Usec[u]] > 0 && rows > 0;
        }
        for (int v = 0; v < nV; ++v) {
                const int64_t cols = size3(sV, Vq[v], 0, 3);
                vsz += cols * kept[Vsec[v]];
                nVo += kept[Vsec[v]] > 0 && cols > 0;
        }
        double* Ut = malloc(sizeof(double) * (usz + 1));
        double* Vt = malloc(sizeof(double) * (vsz + 1));
        int64_t* Uqo = malloc(sizeof(int64_t) * ((size_t) nUo * nincl + 1));
        int64_t* Ubb = malloc(sizeof(int64_t) * (nUo + 2));
        int64_t* Vqo = malloc(sizeof(int64_t) * (nVo + 1));
        int64_t* Vbb = malloc(sizeof(int64_t) * (nVo + 2));
        const t3ns_symsecs_desc* sC = d->ss[id_csite];
        int od[3] = {sC[0].nrSecs, sC[1].nrSecs, sC[2].nrSecs};
        int nd[3] = {od[0], od[1], od[2]};
        nd[id_cbond] = cnt;
        Ubb[0] = 0;
        int bu = 0;
        for (int u = 0; u < nU; ++u) {
                const int s = Usec[u];
                const int64_t* q = Uq + (size_t) u * nincl;
                int64_t t0 = 1, t2 = 1;
                for (int i = 0; i < nincl; ++i) {
                        if (i < csU) t0 *= size3(d->ss[incl[i]], q[i], 0, 3);
                        else if (i > csU) t2 *= size3(d->ss[incl[i]], q[i], 0, 3);
                        else {
                                t0 *= size3(d->ss[incl[i]], q[i], 0, id_cbond);
                                t2 *= size3(d->ss[incl[i]], q[i], id_cbond + 1, 3);
                        }
                }
                const int64_t n = t0 * t2 * kept[s];
                if (n == 0) continue;
                double* dst = Ut + Ubb[bu];
                for (int64_t i2 = 0; i2 < t2; ++i2)
                        for (int t = 0; t < kept[s]; ++t)
                                for (int64_t i0 = 0; i0 < t0; ++i0)
                                        dst[i0 + t * t0 + i2 * t0 * kept[s]] =
                                                US[s][Mrow[u] + i0 + i2 * t0 + (size_t) t * M[s]];
                for (int i = 0; i < nincl; ++i) {
                        int64_t v = q[i];
                        if (i == csU) {
                                int id[3];
                                idx3(id, v, sC);
                                id[id_cbond] = newid[id[id_cbond]];
                                v = id[0] + (int64_t) id[1] * nd[0] + (int64_t) id[2] * nd[0] * nd[1];
                        }
                        Uqo[(size_t) bu * nincl + i] = v;
                }
                Ubb[bu + 1] = Ubb[bu] + n;
                ++bu;
        }
        int odv[3] = {sV[0].nrSecs, sV[1].nrSecs, sV[2].nrSecs};
        int ndv[3] = {odv[0], odv[1], odv[2]};
        ndv[id_bond] = cnt;
        Vbb[0] = 0;
        int bv = 0;
        for (int v = 0; v < nV; ++v) {
                const int s = Vsec[v];
                const int64_t t0 = size3(sV, Vq[v], 0, id_bond), t2 = size3(sV, Vq[v], id_bond + 1, 3);
                const int64_t n = t0 * t2 * kept[s];
                if (n == 0) continue;
                double* dst = Vt + Vbb[bv];
                const int k = nsing[s];
                for (int64_t i2 = 0; i2 < t2; ++i2)
                        for (int t = 0; t < kept[s]; ++t)
                                for (int64_t i0 = 0; i0 < t0; ++i0)
                                        dst[i0 + t * t0 + i2 * t0 * kept[s]] =
                                                VT[s][t + (Ncol[v] + i0 + i2 * t0) * k];
                int id[3];
                idx3(id, Vq[v], sV);
                id[id_bond] = newid[id[id_bond]];
                Vqo[bv] = id[0] + (int64_t) id[1] * ndv[0] + (int64_t) id[2] * ndv[0] * ndv[1];
                Vbb[bv + 1] = Vbb[bv] + n;
                ++bv;
        }
        double nrm = 0;
        for (int64_t i = 0; i < Ubb[bu]; ++i) nrm += Ut[i] * Ut[i];
        nrm = sqrt(nrm);
        for (int64_t i = 0; i < Ubb[bu]; ++i) Ut[i] /= nrm;

        /* output */
        memset(out, 0, sizeof(*out));
        out->U_nrsites = nincl;
        out->U_nrblocks = bu;
        out->U_qnumbers = Uqo;
        out->U_beginblock = Ubb;
        out->U_tel = Ut;
        out->V_nrblocks = bv;
        out->V_qnumbers = Vqo;
        out->V_beginblock = Vbb;
        out->V_tel = Vt;
        int* birr = malloc(sizeof(int) * ((size_t) cnt * NS + 1));
        int* bdim = malloc(sizeof(int) * (cnt + 1));
        for (int s = 0; s < nsec; ++s)
                if (newid[s] >= 0) {
                        memcpy(birr + (size_t) newid[s] * NS, cut->irreps + (size_t) s * NS, sizeof(int) * NS);
                        bdim[newid[s]] = kept[s];
                }
        out->bond_nrSecs = cnt;
        out->bond_irreps = birr;
        out->bond_dims = bdim;
        out->nrsectors = nsec;
        out->nsing = nsing;
        out->kept = kept;
        double* sing = malloc(sizeof(double) * (total + 1));
        c = 0;
        for (int s = 0; s < nsec; ++s)
                for (int t = 0; t < nsing[s]; ++t) sing[c++] = S[s][t];
        out->sing = sing;
        out->norm[0] = norm0;
        out->norm[1] = norm1;
        out->entropy[0] = ent0;
        out->entropy[1] = ent1;
        out->norm_U = nrm;
        for (int s = 0; s < nsec; ++s) {
                free(mat[s]);
                free(S[s]);
                free(US[s]);
                free(VT[s]);
        }
        free(mat); free(S); free(US); free(VT);
        free(Uq); free(Vq); free(Mrow); free(Ncol); free(Usec); free(Vsec); free(M); free(N);
        free(mult); free(all); free(orig); free(order); free(newid); free(ub); free(vb);
        return 0;
}

void oracle_split_free(t3ns_split_view* v)
{
        free((void*) v->U_qnumbers);
        free((void*) v->U_beginblock);
        free((void*) v->U_tel);
        free((void*) v->V_qnumbers);
        free((void*) v->V_beginblock);
        free((void*) v->V_tel);
        free((void*) v->bond_irreps);
        free((void*) v->bond_dims);
        free((void*) v->nsing);
        free((void*) v->kept);
        free((void*) v->sing);
        memset(v, 0, sizeof(*v));
}

/* ------------------------------------------------------------------------
 * qr_step (src/siteTensor_decompose.c:1587-1630): qr (313-355) with
 * init_qrdata / makestart / make_Q (77-196), QR_copy_fromto_mem (214-253),
 * qrblocks (255-311; dgeqrf and dorgqr restated below as unblocked
 * Householder QR with LAPACK's reflector convention), multiplyR with makeB
 * (377-490), R_svd (550-574) for the entropy.
 * ---------------------------------------------------------------------- */

/* dgeqr2 + dorg2r: A (m x n) -> R (k x n upper), Q (m x k), k = min(m, n) */
static void householder(double* A, int m, int n, double* R, double* Q)
{
        const int k = m < n ? m : n;
        double* tau = calloc(k + 1, sizeof(double));
        for (int j = 0; j < k; ++j) {
                double* col = A + (size_t) j * m;
                double xn = 0;
                for (int i = j + 1; i < m; ++i) xn += col[i] * col[i];
                if (xn == 0) {
                        tau[j] = 0;
                } else {
                        const double alpha = col[j];
                        const double nrm = sqrt(alpha * alpha + xn);
                        const double beta = alpha >= 0 ? -nrm : nrm;
                        tau[j] = (beta - alpha) / beta;
                        const double sc = 1 / (alpha - beta);
                        for (int i = j + 1; i < m; ++i) col[i] *= sc;
                        col[j] = beta;
                }
                for (int c = j + 1; c < n && tau[j] != 0; ++c) {
                        double* x = A + (size_t) c * m;
                        double w = x[j];
                        for (int i = j + 1; i < m; ++i) w += col[i] * x[i];
                        x[j] -= tau[j] * w;
                        for (int i = j + 1; i < m; ++i) x[i] -= tau[j] * col[i] * w;
                }
        }
        for (int j = 0; j < n; ++j)
                for (int i = 0; i < k; ++i) R[i + (size_t) j * k] = i <= j ? A[i + (size_t) j * m] : 0;
        for (int j = 0; j < k; ++j)
                for (int i = 0; i < m; ++i) Q[i + (size_t) j * m] = i == j;
        for (int jj = k - 1; jj >= 0; --jj) {
                if (tau[jj] == 0) continue;
                const double* v = A + (size_t) jj * m;
                for (int c = jj; c < k; ++c) {
                        double* x = Q + (size_t) c * m;
                        double w = x[jj];
                        for (int i = jj + 1; i < m; ++i) w += v[i] * x[i];
                        x[jj] -= tau[jj] * w;
                        for (int i = jj + 1; i < m; ++i) x[i] -= tau[jj] * v[i] * w;
                }
        }
        free(tau);
}

int oracle_qr_step(const t3ns_qr_desc* d, t3ns_qr_view* out)
{
        int oc_id = -1, o_id = -1;
        for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                        if (d->ssA[i].bond == d->ssC[j].bond) {
                                oc_id = i;
                                o_id = j;
                        }
        if (oc_id < 0) return -1;
        const t3ns_symsecs_desc* ss = d->ssA;
        const int nsec = ss[oc_id].nrSecs;
        /* blocks of each sector in block order (stable sort of the bond index) */
        int* sec_of = malloc(sizeof(int) * (d->A_nrblocks + 1));
        int64_t* before = malloc(sizeof(int64_t) * (d->A_nrblocks + 1));
        int64_t* after = malloc(sizeof(int64_t) * (d->A_nrblocks + 1));
        int64_t* row0 = malloc(sizeof(int64_t) * (d->A_nrblocks + 1));
        int64_t* M = calloc(nsec + 1, sizeof(int64_t));
        for (int b = 0; b < d->A_nrblocks; ++b) {
                int id[3];
                idx3(id, d->A_qnumbers[b], ss);
                sec_of[b] = id[oc_id];
                before[b] = after[b] = 1;
                for (int i = 0; i < oc_id; ++i) before[b] *= ss[i].dims[id[i]];
                for (int i = oc_id + 1; i < 3; ++i) after[b] *= ss[i].dims[id[i]];
                row0[b] = M[id[oc_id]];
                if (d->A_beginblock[b + 1] > d->A_beginblock[b]) M[id[oc_id]] += before[b] * after[b];
        }
        int* K = calloc(nsec + 1, sizeof(int));
        double** Rm = calloc(nsec + 1, sizeof(double*));
        double** Qm = calloc(nsec + 1, sizeof(double*));
        for (int s = 0; s < nsec; ++s) {
                const int64_t N = M[s] ? ss[oc_id].dims[s] : 0;
                K[s] = (int) (M[s] < N ? M[s] : N);
                if (!K[s]) continue;
                double* mat = calloc((size_t) M[s] * N + 1, sizeof(double));
                for (int b = 0; b < d->A_nrblocks; ++b) {
                        if (sec_of[b] != s || d->A_beginblock[b + 1] == d->A_beginblock[b]) continue;
                        const double* src = d->A_tel + d->A_beginblock[b];
                        for (int64_t o = 0; o < after[b]; ++o)
                                for (int64_t n = 0; n < N; ++n)
                                        for (int64_t m = 0; m < before[b]; ++m)
                                                mat[row0[b] + m + o * before[b] + n * M[s]] =
                                                        src[m + n * before[b] + o * before[b] * N];
                }
                Rm[s] = calloc((size_t) K[s] * N + 1, sizeof(double));
                Qm[s] = calloc((size_t) M[s] * K[s] + 1, sizeof(double));
                householder(mat, (int) M[s], (int) N, Rm[s], Qm[s]);
                free(mat);
        }
        /* Q blocks (QR_copy_fromto_mem FROM_MEMORY) */
        int64_t* Qbb = calloc((size_t) d->A_nrblocks + 2, sizeof(int64_t));
        for (int b = 0; b < d->A_nrblocks; ++b) {
                const int64_t sz = d->A_beginblock[b + 1] == d->A_beginblock[b] ? 0 : before[b] * K[sec_of[b]] * after[b];
                Qbb[b + 1] = Qbb[b] + sz;
        }
        double* Qt = calloc((size_t) Qbb[d->A_nrblocks] + 1, sizeof(double));
        for (int b = 0; b < d->A_nrblocks; ++b) {
                const int s = sec_of[b];
                if (Qbb[b + 1] == Qbb[b]) continue;
                for (int64_t o = 0; o < after[b]; ++o)
                        for (int64_t n = 0; n < K[s]; ++n)
                                for (int64_t m = 0; m < before[b]; ++m)
                                        Qt[Qbb[b] + m + n * before[b] + o * before[b] * K[s]] =
                                                Qm[s][row0[b] + m + o * before[b] + n * M[s]];
        }
        /* multiplyR (bondR = 1): B = C contracted over its bond o_id with R */
        const t3ns_symsecs_desc* sc = d->ssC;
        int64_t* Bbb = calloc((size_t) d->C_nrblocks + 2, sizeof(int64_t));
        for (int b = 0; b < d->C_nrblocks; ++b) {
                int id[3];
                idx3(id, d->C_qnumbers[b], sc);
                const int64_t sz = d->C_beginblock[b + 1] - d->C_beginblock[b];
                const int s = id[o_id];
                Bbb[b + 1] = Bbb[b] + (K[s] == 0 || sz == 0 ? 0 : sz / sc[o_id].dims[s] * K[s]);
        }
        double* Bt = calloc((size_t) Bbb[d->C_nrblocks] + 1, sizeof(double));
        for (int b = 0; b < d->C_nrblocks; ++b) {
                if (Bbb[b + 1] == Bbb[b]) continue;
                int id[3];
                idx3(id, d->C_qnumbers[b], sc);
                const int s = id[o_id];
                const int N = sc[o_id].dims[s];
                const double* A = d->C_tel + d->C_beginblock[b];
                const double* R = Rm[s];
                double* Bo = Bt + Bbb[b];
                /* the reference's contractinfo (multiplyR 443-484) */
                const int dim1A = sc[0].dims[id[0]] * (o_id == 1 ? 1 : sc[1].dims[id[1]]);
                const int dim2A = sc[2].dims[id[2]] * (o_id == 1 ? 1 : sc[1].dims[id[1]]);
                const int bondA = o_id, bondR = 1;
                const int trB = !((bondR == 0) == (bondA != 0));
                const int Mx = bondA == 0 ? K[s] : dim1A;
                const int Nx = bondA == 0 ? dim2A : K[s];
                const int Kx = N;
                const int L = bondA == 1 ? dim2A : 1;
                if (bondA == 0) {
                        /* C = op(R) A */
                        oracle_dgemm_batched(trB, 0, Mx, Nx, Kx, L, 1.0, R, trB ? Kx : Mx, 0, A, Kx,
                                             (int64_t) Mx * Kx, 0.0, Bo, Mx, (int64_t) Mx * Nx);
                } else {
                        /* C = A op(R), batched over the last leg for the middle bond */
                        oracle_dgemm_batched(0, trB, Mx, Nx, Kx, L, 1.0, A, Mx, (int64_t) Mx * Kx, R,
                                             trB ? Nx : Kx, 0, 0.0, Bo, Mx, (int64_t) Mx * Nx);
                }
        }
        /* R_svd and the entropy */
        int* nsing = calloc(nsec + 1, sizeof(int));
        int total = 0;
        for (int s = 0; s < nsec; ++s) total += K[s];
        double* sing = calloc(total + 1, sizeof(double));
        double ent = 0;
        int c = 0;
        for (int s = 0; s < nsec; ++s) {
                nsing[s] = K[s];
                if (!K[s]) continue;
                const int N = ss[oc_id].dims[s];
                double* US = malloc(sizeof(double) * ((size_t) K[s] * K[s] + 1));
                double* VT = malloc(sizeof(double) * ((size_t) K[s] * N + 1));
                svd(Rm[s], K[s], N, sing + c, US, VT);
                const int mult = multiplicity_of(ss[oc_id].irreps + (size_t) s * NS, d->sgs, d->nrSyms);
                double e = 0;
                for (int t = 0; t < K[s]; ++t) e += vn_entropy(sing[c + t] / sqrt((double) mult));
                ent += mult * e;
                c += K[s];
                free(US);
                free(VT);
        }
        memset(out, 0, sizeof(*out));
        out->nrblocks_Q = d->A_nrblocks;
        out->Q_beginblock = Qbb;
        out->Q_tel = Qt;
        out->nrblocks_B = d->C_nrblocks;
        out->B_beginblock = Bbb;
        out->B_tel = Bt;
        out->nrsectors = nsec;
        out->dims = K;
        out->nsing = nsing;
        out->sing = sing;
        out->entropy = ent;
        for (int s = 0; s < nsec; ++s) {
                free(Rm[s]);
                free(Qm[s]);
        }
        free(Rm); free(Qm); free(sec_of); free(before); free(after); free(row0); free(M);
        return 0;
}

void oracle_qr_free(t3ns_qr_view* v)
{
        free((void*) v->Q_beginblock);
        free((void*) v->Q_tel);
        free((void*) v->B_beginblock);
        free((void*) v->B_tel);
        free((void*) v->dims);
        free((void*) v->nsing);
        free((void*) v->sing);
        memset(v, 0, sizeof(*v));
}
