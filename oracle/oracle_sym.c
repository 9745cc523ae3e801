/*
 * ORACLE -- test infrastructure.  Symmetry prefactors and Wigner symbols.
 * Restates src/symmetries.c:158-332, src/symmetry_z2.c:54-228,
 * src/symmetry_su2.c:26-120 and the Racah formula of src/Wigner.c:245-330.
 * Irrep arrays are flat int arrays with T3NS_MAX_SYMMETRIES entries per irrep.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define NS T3NS_MAX_SYMMETRIES
enum { SG_Z2 = 0, SG_U1 = 1, SG_SU2 = 2 };

static long double g_fact[401];
static int g_fact_ready = 0;

static void init_fact(void)
{
        if (g_fact_ready) return;
        g_fact[0] = 1.0L;
        for (int i = 1; i <= 400; ++i) g_fact[i] = g_fact[i - 1] * (long double) i;
        g_fact_ready = 1;
}

/* triangle rule on doubled spins, src/Wigner.c:230-235 */
static int tri_bad(int a, int b, int c)
{
        return ((a + b + c) % 2 != 0) || (a + b < c) || (abs(a - b) > c);
}

static long double tri_coef(int a, int b, int c)
{
        return g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] *
                g_fact[(-a + b + c) / 2] / g_fact[(a + b + c) / 2 + 1];
}

/* src/Wigner.c:245-296: {a b c; d e f} by Racah's single sum */
double oracle_wigner6j(int a, int b, int c, int d, int e, int f)
{
        init_fact();
        if (tri_bad(a, b, c) || tri_bad(d, e, c) || tri_bad(a, e, f) || tri_bad(d, b, f))
                return 0.0;
        const int al[4] = {(a + b + c) / 2, (d + e + c) / 2, (a + e + f) / 2, (d + b + f) / 2};
        const int be[3] = {(a + b + d + e) / 2, (a + c + d + f) / 2, (b + c + e + f) / 2};
        int lo = al[0], hi = be[0];
        for (int i = 1; i < 4; ++i) if (al[i] > lo) lo = al[i];
        for (int i = 1; i < 3; ++i) if (be[i] < hi) hi = be[i];
        if (hi < lo) return 0.0;
        long double s = 0.0L;
        for (int k = lo; k <= hi; ++k) {
                long double den = 1.0L;
                for (int i = 0; i < 4; ++i) den *= g_fact[k - al[i]];
                for (int i = 0; i < 3; ++i) den *= g_fact[be[i] - k];
                const long double t = g_fact[k + 1] / den;
                s += (k % 2) ? -t : t;
        }
        const long double n = sqrtl(tri_coef(a, b, c) * tri_coef(d, e, c) *
                                    tri_coef(a, e, f) * tri_coef(d, b, f));
        return (double) (s * n);
}

/* src/Wigner.c:298-330: 9j as a sum over x of three 6j symbols */
double oracle_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i)
{
        if (tri_bad(a, b, c) || tri_bad(d, e, f) || tri_bad(g, h, i) ||
            tri_bad(a, d, g) || tri_bad(b, e, h) || tri_bad(c, f, i))
                return 0.0;
        int lo = abs(b - f);
        if (abs(h - d) > lo) lo = abs(h - d);
        if (abs(a - i) > lo) lo = abs(a - i);
        int hi = a + i;
        if (h + d < hi) hi = h + d;
        if (b + f < hi) hi = b + f;
        double v = 0.0;
        for (int x = lo; x <= hi; x += 2) {
                v += (x + 1) * oracle_wigner6j(a, b, c, f, i, x) *
                        oracle_wigner6j(d, e, f, b, x, h) *
                        oracle_wigner6j(g, h, i, x, a, d);
        }
        return (lo % 2 == 0) ? v : -v;
}

static double br(int twoj) { return sqrt(twoj + 1.0); }
static double sgn(int odd) { return odd ? -1.0 : 1.0; }

/* irr layout [j][k][sym] -> irr[(j*3 + k)*NS + sym] */
#define IR(j, k, y) irr[((j) * 3 + (k)) * NS + (y)]

double oracle_prefactor_add_P_operator(const int* irr, int isleft, const int* sgs, int nrsy)
{
        /* src/symmetries.c:280-299, src/symmetry_z2.c:161-209 */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] != SG_Z2 || isleft) continue;
                p *= sgn((IR(1, 2, y) + IR(1, 0, y)) % 2);
        }
        return p;
}

double oracle_prefactor_combine_MPOs(const int* irr, const int* mpo, const int* sgs, int nrsy,
                                     int isdmrg, int e)
{
        /* src/symmetries.c:301-332, src/symmetry_z2.c:211-228,
         * src/symmetry_su2.c:95-108 */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                const int m0 = mpo[0 * NS + y], m1 = mpo[1 * NS + y], m2 = mpo[2 * NS + y];
                if (sgs[y] == SG_Z2) {
                        if (isdmrg)
                                p *= sgn((IR(1, e, y) + IR(0, e, y) * m1) % 2);
                        else
                                p *= sgn((IR(1, 1, y) * m0 + IR(1, 2, y) + IR(0, 2, y) * m2) % 2);
                } else if (sgs[y] == SG_SU2) {
                        if (isdmrg) {
                                p *= sgn((IR(0, e, y) + IR(1, e, y) + m1) % 4) /
                                        (br(IR(0, e, y)) * br(IR(1, e, y)));
                        } else {
                                p *= sgn((IR(0, 2, y) + IR(1, 2, y) + m2) % 4) * br(m2) *
                                        oracle_wigner9j(IR(0, 0, y), IR(1, 0, y), m0,
                                                        IR(0, 1, y), IR(1, 1, y), m1,
                                                        IR(0, 2, y), IR(1, 2, y), m2);
                        }
                }
        }
        return p;
}

double oracle_prefactor_pAppend(const int* irr, int is_left, const int* sgs, int nrsy)
{
        /* src/symmetries.c:158-183, src/symmetry_z2.c:54-60,
         * src/symmetry_su2.c:75-93 */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] == SG_Z2) {
                        if (is_left)
                                p *= sgn(IR(1, 1, y) && IR(2, 0, y));
                        else
                                p *= sgn((IR(2, 1, y) * IR(1, 2, y) + IR(2, 1, y)) % 2);
                } else if (sgs[y] == SG_SU2) {
                        if (is_left) {
                                p *= br(IR(0, 2, y)) * br(IR(1, 2, y)) * br(IR(2, 2, y)) *
                                        oracle_wigner9j(IR(0, 0, y), IR(1, 0, y), IR(2, 0, y),
                                                        IR(0, 1, y), IR(1, 1, y), IR(2, 1, y),
                                                        IR(0, 2, y), IR(1, 2, y), IR(2, 2, y));
                        } else {
                                p *= sgn((IR(0, 1, y) + IR(1, 1, y) + IR(2, 1, y)) % 4) *
                                        br(IR(0, 0, y)) * br(IR(1, 0, y)) * br(IR(2, 2, y)) *
                                        oracle_wigner9j(IR(0, 0, y), IR(1, 0, y), IR(2, 0, y),
                                                        IR(0, 2, y), IR(1, 2, y), IR(2, 2, y),
                                                        IR(0, 1, y), IR(1, 1, y), IR(2, 1, y));
                        }
                }
        }
        return p;
}

double oracle_prefactor_adjoint(const int* irr, char c, const int* sgs, int nrsy)
{
        /* src/symmetries.c:185-201, src/symmetry_z2.c:62-84; irr is [3][NS] */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] != SG_Z2) continue;
                switch (c) {
                case 'c': p *= sgn(irr[2 * NS + y]); break;
                case '1': p *= sgn(irr[1 * NS + y]); break;
                case '2': p *= sgn(irr[0 * NS + y]); break;
                default: break;
                }
        }
        return p;
}

double oracle_prefactor_pUpdate(const int* irr, int is_left, const int* sgs, int nrsy)
{
        /* src/symmetries.c:213-235, src/symmetry_z2.c:86-103 */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] != SG_Z2 || is_left) continue;
                p *= sgn((IR(0, 1, y) + IR(1, 1, y)) % 2);
        }
        return p;
}

double oracle_prefactor_mirror_coupling(const int* irr, const int* sgs, int nrsy)
{
        /* src/symmetries.c:237-251; irr is [3][NS] */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] == SG_Z2)
                        p *= sgn((irr[0 * NS + y] + irr[1 * NS + y] * irr[2 * NS + y]) % 2);
                else if (sgs[y] == SG_SU2)
                        p *= br(irr[0 * NS + y]);
        }
        return p;
}

double oracle_prefactor_bUpdate(const int* irr, int u, const int* sgs, int nrsy)
{
        /* src/symmetries.c:253-278, src/symmetry_z2.c:105-153,
         * src/symmetry_su2.c:110-120 */
        double p = 1;
        for (int y = 0; y < nrsy; ++y) {
                if (sgs[y] == SG_Z2) {
                        if (u == 0) p *= sgn((IR(2, 1, y) * IR(1, 2, y) + IR(1, 0, y)) % 2);
                        else if (u == 1) p *= sgn((IR(0, 1, y) * IR(2, 2, y) + IR(0, 0, y)) % 2);
                        else p *= sgn(IR(1, 1, y) && IR(0, 2, y));
                } else if (sgs[y] == SG_SU2) {
                        const int s = (IR(2, 0, y) + IR(2, 1, y) + IR(2, 2, y) +
                                       IR(u, 0, y) + IR(u, 1, y) + IR(u, 2, y)) % 4;
                        p *= sgn(s) * br(IR(u, 0, y)) * br(IR(u, 1, y)) * br(IR(2, 2, y)) *
                                oracle_wigner9j(IR(0, 0, y), IR(0, 1, y), IR(0, 2, y),
                                                IR(1, 0, y), IR(1, 1, y), IR(1, 2, y),
                                                IR(2, 0, y), IR(2, 1, y), IR(2, 2, y));
                }
        }
        return p;
}

/* The reference's do_contract calls cblas_dgemm (src/sparseblocks.c:144,
 * 153).  oracle_use_blas() points the oracle at an installed CBLAS dgemm
 * (scipy's OpenBLAS, symbol prefix scipy_) so the CPU baseline runs the
 * same kind of GEMM; without it the loop nest below serves.  CBLAS
 * constants: column major 102, no transpose 111, transpose 112. */
typedef void (*cblas_dgemm_fn)(int, int, int, int, int, int, double, const double*, int,
                               const double*, int, double, double*, int);
static cblas_dgemm_fn blas_dgemm = NULL;

int oracle_use_blas(const char* path)
{
        void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!h) return -1;
        cblas_dgemm_fn f = (cblas_dgemm_fn) dlsym(h, "scipy_cblas_dgemm");
        void (*threads)(int) = (void (*)(int)) dlsym(h, "scipy_openblas_set_num_threads");
        if (!f || !threads) return -1;
        threads(1);  /* one BLAS thread inside each OpenMP thread, as the reference runs */
        blas_dgemm = f;
        return 0;
}

int oracle_blas_active(void) { return blas_dgemm != NULL; }

/* src/sparseblocks.c:133-158: C_l = beta C_l + alpha op(A_l) op(B_l), l < L */
void oracle_dgemm_batched(int ta, int tb, int M, int N, int K, int L, double alpha,
                          const double* A, int lda, int64_t sA, const double* B, int ldb,
                          int64_t sB, double beta, double* C, int ldc, int64_t sC)
{
        /* BLAS only where its argument rules hold (leading dimensions at
         * least the stored rows); it would refuse the call otherwise */
        if (blas_dgemm && M > 0 && N > 0 && K > 0 && lda >= (ta ? K : M) &&
            ldb >= (tb ? N : K) && ldc >= M) {
                for (int l = 0; l < L; ++l)
                        blas_dgemm(102, ta ? 112 : 111, tb ? 112 : 111, M, N, K, alpha, A + l * sA,
                                   lda, B + l * sB, ldb, beta, C + l * sC, ldc);
                return;
        }
        for (int l = 0; l < L; ++l) {
                const double* Al = A + l * sA;
                const double* Bl = B + l * sB;
                double* Cl = C + l * sC;
                for (int j = 0; j < N; ++j) {
                        double* c = Cl + (int64_t) j * ldc;
                        if (beta == 0.0) {
                                for (int i = 0; i < M; ++i) c[i] = 0.0;
                        } else if (beta != 1.0) {
                                for (int i = 0; i < M; ++i) c[i] *= beta;
                        }
                        for (int k = 0; k < K; ++k) {
                                const double b = alpha * (tb ? Bl[j + (int64_t) k * ldb]
                                                             : Bl[k + (int64_t) j * ldb]);
                                if (!ta) {
                                        const double* a = Al + (int64_t) k * lda;
                                        for (int i = 0; i < M; ++i) c[i] += a[i] * b;
                                } else {
                                        for (int i = 0; i < M; ++i)
                                                c[i] += Al[k + (int64_t) i * lda] * b;
                                }
                        }
                }
        }
}
