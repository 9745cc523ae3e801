// SVD split on the device (src/siteTensor_decompose.c:587-1444).
#include "decompose_dev.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace t3 {

namespace {

constexpr int kJThreads = 256;

// Pair (p, q) of round r of the circle schedule over np (even) columns.
__device__ __forceinline__ void round_pair(int np, int r, int k, int* p, int* q) {
    const int m = np - 1;
    if (k == 0) {
        *p = 0;
        *q = 1 + r % m;
    } else {
        *p = 1 + (r + k) % m;
        *q = 1 + (r - k + m) % m;
    }
}

struct JItem {
    int mat, k;  // matrix, pair slot in the round
};

// One Jacobi round: every work item rotates one column pair of one matrix.
__global__ void __launch_bounds__(kJThreads)
    jacobi_round_kernel(const JacobiMat* __restrict__ mats, const JItem* __restrict__ items, int nitems,
                        int r, double tol, double* arena, int* rotated) {
    __shared__ double red[3][kJThreads];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const JItem w = items[it];
        const JacobiMat M = mats[w.mat];
        const int np = M.n + (M.n & 1);
        if (r >= np - 1) continue;  // uniform over the workgroup
        int p, q;
        round_pair(np, r, w.k, &p, &q);
        if (p >= M.n || q >= M.n) continue;
        if (p > q) {
            const int t = p;
            p = q;
            q = t;
        }
        double* bp = arena + M.b + (int64_t)p * M.m;
        double* bq = arena + M.b + (int64_t)q * M.m;
        double a = 0, b = 0, g = 0;
        for (int i = threadIdx.x; i < M.m; i += kJThreads) {
            const double x = bp[i], y = bq[i];
            a = fma(x, x, a);
            b = fma(y, y, b);
            g = fma(x, y, g);
        }
        red[0][threadIdx.x] = a;
        red[1][threadIdx.x] = b;
        red[2][threadIdx.x] = g;
        __syncthreads();
        for (int st = kJThreads / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st)
                for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
            __syncthreads();
        }
        a = red[0][0];
        b = red[1][0];
        g = red[2][0];
        __syncthreads();
        // the dot product carries a rounding error growing as sqrt(rows)
        if (g != 0 && fabs(g) > tol * sqrt((double)M.m) * sqrt(a * b)) {
            const double zeta = (b - a) / (2 * g);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
            const double c = 1 / sqrt(1 + t * t), s = c * t;
            for (int i = threadIdx.x; i < M.m; i += kJThreads) {
                const double x = bp[i], y = bq[i];
                bp[i] = c * x - s * y;
                bq[i] = s * x + c * y;
            }
            double* vp = arena + M.v + (int64_t)p * M.n;
            double* vq = arena + M.v + (int64_t)q * M.n;
            for (int i = threadIdx.x; i < M.n; i += kJThreads) {
                const double x = vp[i], y = vq[i];
                vp[i] = c * x - s * y;
                vq[i] = s * x + c * y;
            }
            if (threadIdx.x == 0) atomicOr(rotated, 1);
        }
    }
}

__global__ void identity_kernel(const JacobiMat* __restrict__ mats, int nm, double* arena) {
    for (int m = blockIdx.x; m < nm; m += gridDim.x) {
        const JacobiMat M = mats[m];
        const int64_t nn = (int64_t)M.n * M.n;
        for (int64_t i = threadIdx.x; i < nn; i += blockDim.x)
            arena[M.v + i] = (i % (M.n + 1) == 0) ? 1.0 : 0.0;
    }
}

__global__ void colnorm_kernel(const JacobiMat* __restrict__ mats, int nm, const double* arena,
                               const int64_t* __restrict__ out_off, double* out) {
    __shared__ double red[kJThreads];
    for (int m = blockIdx.x; m < nm; m += gridDim.x) {
        const JacobiMat M = mats[m];
        for (int j = 0; j < M.n; ++j) {
            const double* col = arena + M.b + (int64_t)j * M.m;
            double a = 0;
            for (int i = threadIdx.x; i < M.m; i += kJThreads) a = fma(col[i], col[i], a);
            red[threadIdx.x] = a;
            __syncthreads();
            for (int st = kJThreads / 2; st > 0; st >>= 1) {
                if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
                __syncthreads();
            }
            if (threadIdx.x == 0) out[out_off[m] + j] = sqrt(red[0]);
            __syncthreads();
        }
    }
}

// A blocks into sector matrices: element (a, v, c) of a block with dims
// (d0, d1, d2) goes to row row0 + a + c * d0, column col0 + v.
struct Gather {
    int64_t src, mat;  // A offset, matrix offset
    int64_t row0, col0, ld;
    int64_t d0, d1, d2;
    int transposed;    // store the transpose (ld is then the column count)
    int pad;
};

__global__ void gather_kernel(const Gather* __restrict__ g, int n, const double* __restrict__ A,
                              double* arena) {
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const Gather G = g[b];
        const int64_t tot = G.d0 * G.d1 * G.d2;
        for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
            const int64_t a = e % G.d0, v = (e / G.d0) % G.d1, c = e / (G.d0 * G.d1);
            const int64_t row = G.row0 + a + c * G.d0, col = G.col0 + v;
            const int64_t dst = G.transposed ? col + row * G.ld : row + col * G.ld;
            arena[G.mat + dst] = A[G.src + e];
        }
    }
}

// Kept singular vectors into a tensor block with dims (t0, kept, t2):
// element (i0, t, i2) = scale[t] * src[row0 + i0 + i2 * t0 + col(t) * ld]
// with col(t) = perm[t] (the t-th largest singular value).
struct Scatter {
    int64_t dst, src, row0, ld;
    int64_t t0, t2;
    int kept, sector;
};

__global__ void scatter_kernel(const Scatter* __restrict__ sc, int n, const int* __restrict__ perm,
                               const int64_t* __restrict__ perm_off, const double* __restrict__ scale,
                               const double* arena, double* __restrict__ out) {
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const Scatter S = sc[b];
        const int64_t tot = S.t0 * S.kept * S.t2;
        for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
            const int64_t i0 = e % S.t0, t = (e / S.t0) % S.kept, i2 = e / (S.t0 * S.kept);
            const int64_t po = perm_off[S.sector] + t;
            const int col = perm[po];
            out[S.dst + e] = scale[po] * arena[S.src + S.row0 + i0 + i2 * S.t0 + (int64_t)col * S.ld];
        }
    }
}

__global__ void sumsq_kernel(const double* __restrict__ x, int64_t n, double* out) {
    __shared__ double red[1024];
    double a = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a = fma(x[i], x[i], a);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void scale_kernel(double* x, int64_t n, double f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= f;
}

template <typename T>
__global__ void zero_kernel(T* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

int gridn(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(16384, n)); }

int64_t block_prod(const Symsecs* ss, QN qn, int from, int to) {
    const QN n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
    const int id[3] = {(int)(qn % n0), (int)((qn / n0) % n1), (int)(qn / (n0 * n1))};
    int64_t r = 1;
    for (int k = from; k < to; ++k) r *= ss[k].dims[id[k]];
    return r;
}

}  // namespace

int jacobi_svd(double* arena, const std::vector<JacobiMat>& mats, hipStream_t s, int max_sweeps,
               double tol) {
    if (mats.empty()) return 0;
    DevBuf<JacobiMat> dm;
    dm.upload(mats, s);
    std::vector<JItem> items;
    int maxround = 0;
    for (size_t m = 0; m < mats.size(); ++m) {
        if (mats[m].m < mats[m].n) throw std::invalid_argument("jacobi: rows < columns");
        const int np = mats[m].n + (mats[m].n & 1);
        for (int k = 0; k < np / 2; ++k) items.push_back(JItem{(int)m, k});
        maxround = std::max(maxround, np - 1);
    }
    DevBuf<JItem> di;
    di.upload(items, s);
    DevBuf<int> flag(1);
    hipLaunchKernelGGL(identity_kernel, dim3(gridn((int64_t)mats.size())), dim3(256), 0, s, dm.get(),
                       (int)mats.size(), arena);
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
        hipLaunchKernelGGL(zero_kernel<int>, dim3(1), dim3(64), 0, s, flag.get(), (int64_t)1);
        for (int r = 0; r < maxround; ++r)
            hipLaunchKernelGGL(jacobi_round_kernel, dim3(gridn((int64_t)items.size())), dim3(kJThreads), 0,
                               s, dm.get(), di.get(), (int)items.size(), r, tol, arena, flag.get());
        int h = 0;
        T3_HIP_CHECK(hipMemcpyAsync(&h, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        if (!h) break;
    }
    T3_HIP_CHECK(hipGetLastError());
    return sweep + 1;
}

SplitResult run_split(const SplitInput& in, const SplitPlan& P, const std::vector<double>& A,
                      hipStream_t s, RenormWork* work) {
    SplitResult R;
    const int nsec = P.nrSss;
    // per sector: B = the sector matrix with rows >= columns (its transpose
    // when it has more columns), V its right rotations
    std::vector<JacobiMat> mats;
    std::vector<int> mat_of(nsec, -1);
    std::vector<char> trans(nsec, 0);
    int64_t top = 0;
    double flops = 0;
    for (int ss = 0; ss < nsec; ++ss) {
        const SplitSector& S = P.sec[ss];
        if (S.dimS == 0) continue;
        JacobiMat J;
        trans[ss] = S.M < S.N;
        J.m = (int)(trans[ss] ? S.N : S.M);
        J.n = (int)(trans[ss] ? S.M : S.N);
        J.b = top;
        top += (int64_t)J.m * J.n;
        J.v = top;
        top += (int64_t)J.n * J.n;
        mat_of[ss] = (int)mats.size();
        mats.push_back(J);
        flops += 6.0 * J.m * (double)J.n * J.n;  // per sweep, rotations of B and V
    }
    std::vector<Gather> gs;
    for (size_t b = 0; b < P.ablocks.size(); ++b) {
        const SplitPlan::ABlock& ab = P.ablocks[b];
        const int mi = mat_of[ab.sector];
        if (mi < 0) continue;
        const SplitSector& S = P.sec[ab.sector];
        Gather g;
        g.src = in.beginblock[b];
        g.mat = mats[mi].b;
        g.row0 = ab.row0;
        g.col0 = ab.col0;
        g.transposed = trans[ab.sector];
        g.ld = trans[ab.sector] ? S.N : S.M;
        g.d0 = ab.d[0];
        g.d1 = ab.d[1];
        g.d2 = ab.d[2];
        g.pad = 0;
        gs.push_back(g);
    }
    DevBuf<double> dA, arena(std::max<int64_t>(1, top));
    dA.upload(A, s);
    if (A.empty()) dA.alloc(1);
    hipLaunchKernelGGL(zero_kernel<double>, dim3(gridn((top + 255) / 256)), dim3(256), 0, s, arena.get(),
                       top);
    DevBuf<Gather> dg;
    dg.upload(gs, s);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    if (!gs.empty())
        hipLaunchKernelGGL(gather_kernel, dim3(gridn((int64_t)gs.size())), dim3(256), 0, s, dg.get(),
                           (int)gs.size(), dA.get(), arena.get());
    R.sweeps = jacobi_svd(arena.get(), mats, s, 40, 1e-15);
    flops *= R.sweeps;
    // singular values: column norms of the rotated B
    std::vector<int64_t> soff(mats.size() + 1, 0);
    for (size_t m = 0; m < mats.size(); ++m) soff[m + 1] = soff[m] + mats[m].n;
    DevBuf<JacobiMat> dm;
    dm.upload(mats, s);
    DevBuf<int64_t> dso;
    dso.upload(soff, s);
    DevBuf<double> dsig(std::max<int64_t>(1, soff.back()));
    if (!mats.empty())
        hipLaunchKernelGGL(colnorm_kernel, dim3(gridn((int64_t)mats.size())), dim3(kJThreads), 0, s,
                           dm.get(), (int)mats.size(), arena.get(), dso.get(), dsig.get());
    std::vector<double> sig(soff.back());
    if (!sig.empty())
        T3_HIP_CHECK(hipMemcpyAsync(sig.data(), dsig.get(), sig.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));

    // descending order per sector; truncation (selectS)
    R.sing.assign(nsec, {});
    std::vector<std::vector<int>> order(nsec);
    for (int ss = 0; ss < nsec; ++ss) {
        const int mi = mat_of[ss];
        if (mi < 0) continue;
        const int n = mats[mi].n;
        order[ss].resize(n);
        std::iota(order[ss].begin(), order[ss].end(), 0);
        const double* sv = sig.data() + soff[mi];
        std::stable_sort(order[ss].begin(), order[ss].end(), [&](int a, int b) { return sv[a] > sv[b]; });
        for (int j : order[ss]) R.sing[ss].push_back(sv[j]);
    }
    const Symsecs& cut = in.ss[P.id_siteV][P.id_bond];
    std::vector<int> mult(nsec);
    for (int ss = 0; ss < nsec; ++ss) mult[ss] = multiplicity(in.sym, cut.irreps[ss].data());
    R.kept = select_singular_values(R.sing, mult, in.sel, &R.res);

    // U and V structures with the kept dims, before removing empty sectors
    const int nU = (int)(P.U_qn.size() / std::max(1, P.U_nrsites)), nV = (int)P.V_qn.size();
    std::vector<int64_t> ubb(nU + 1, 0), vbb(nV + 1, 0);
    std::vector<int64_t> usz(nU, 0), vsz(nV, 0);
    for (int ss = 0; ss < nsec; ++ss) {
        const SplitSector& S = P.sec[ss];
        for (size_t k = 0; k < S.idU.size(); ++k) usz[S.idU[k]] = (S.Mstart[k + 1] - S.Mstart[k]) * R.kept[ss];
        for (size_t k = 0; k < S.idV.size(); ++k) vsz[S.idV[k]] = (S.Nstart[k + 1] - S.Nstart[k]) * R.kept[ss];
    }
    for (int i = 0; i < nU; ++i) ubb[i + 1] = ubb[i] + usz[i];
    for (int i = 0; i < nV; ++i) vbb[i + 1] = vbb[i] + vsz[i];

    // scatter: U gets the singular vectors times the singular values, V the
    // right singular vectors (SVD_copy_from_mem)
    std::vector<int> perm;
    std::vector<double> uscale, vscale;
    std::vector<int64_t> poff(nsec + 1, 0);
    for (int ss = 0; ss < nsec; ++ss) {
        poff[ss + 1] = poff[ss] + R.kept[ss];
        for (int t = 0; t < R.kept[ss]; ++t) {
            const int j = order[ss][t];
            const double sg = R.sing[ss][t];
            perm.push_back(j);
            // rows >= columns: B holds U * Sigma, V the right vectors;
            // transposed: B holds (right vectors) * Sigma, V the left vectors
            uscale.push_back(trans[ss] ? sg : 1.0);
            vscale.push_back(trans[ss] ? (sg > 0 ? 1.0 / sg : 0.0) : 1.0);
        }
    }
    const Symsecs* sV = in.ss[P.id_siteV];
    std::vector<const Symsecs*> ussU;
    for (int i = 0; i < in.nrsites; ++i)
        if (i != P.id_siteV) ussU.push_back(in.ss[i]);
    const int csiteU = P.id_csite - (P.id_csite > P.id_siteV);
    std::vector<Scatter> scU, scV;
    for (int ss = 0; ss < nsec; ++ss) {
        if (R.kept[ss] == 0) continue;
        const SplitSector& S = P.sec[ss];
        const JacobiMat& J = mats[mat_of[ss]];
        for (size_t k = 0; k < S.idU.size(); ++k) {
            const int u = S.idU[k];
            // dims before / after the cut leg in U's site and leg order
            int64_t t0 = 1, t2 = 1;
            for (int i = 0; i < P.U_nrsites; ++i) {
                const QN q = P.U_qn[(size_t)u * P.U_nrsites + i];
                if (i < csiteU) t0 *= block_prod(ussU[i], q, 0, 3);
                else if (i > csiteU) t2 *= block_prod(ussU[i], q, 0, 3);
                else {
                    t0 *= block_prod(ussU[i], q, 0, P.id_cbond);
                    t2 *= block_prod(ussU[i], q, P.id_cbond + 1, 3);
                }
            }
            Scatter sc;
            sc.dst = ubb[u];
            sc.src = trans[ss] ? J.v : J.b;
            sc.ld = trans[ss] ? J.n : J.m;
            sc.row0 = S.Mstart[k];
            sc.t0 = t0;
            sc.t2 = t2;
            sc.kept = R.kept[ss];
            sc.sector = ss;
            scU.push_back(sc);
        }
        for (size_t k = 0; k < S.idV.size(); ++k) {
            const int v = S.idV[k];
            const QN q = P.V_qn[v];
            Scatter sc;
            sc.dst = vbb[v];
            sc.src = trans[ss] ? J.b : J.v;
            sc.ld = trans[ss] ? J.m : J.n;
            sc.row0 = S.Nstart[k];
            sc.t0 = block_prod(sV, q, 0, P.id_bond);
            sc.t2 = block_prod(sV, q, P.id_bond + 1, 3);
            sc.kept = R.kept[ss];
            sc.sector = ss;
            scV.push_back(sc);
        }
    }
    DevBuf<int> dperm;
    DevBuf<int64_t> dpoff;
    DevBuf<double> dus, dvs, dU(std::max<int64_t>(1, ubb[nU])), dV(std::max<int64_t>(1, vbb[nV]));
    DevBuf<Scatter> dscU, dscV;
    dperm.upload(perm, s);
    dpoff.upload(poff, s);
    dus.upload(uscale, s);
    dvs.upload(vscale, s);
    dscU.upload(scU, s);
    dscV.upload(scV, s);
    if (perm.empty()) {
        dperm.alloc(1);
        dus.alloc(1);
        dvs.alloc(1);
    }
    if (!scU.empty())
        hipLaunchKernelGGL(scatter_kernel, dim3(gridn((int64_t)scU.size())), dim3(256), 0, s, dscU.get(),
                           (int)scU.size(), dperm.get(), dpoff.get(), dus.get(), arena.get(), dU.get());
    if (!scV.empty())
        hipLaunchKernelGGL(scatter_kernel, dim3(gridn((int64_t)scV.size())), dim3(256), 0, s, dscV.get(),
                           (int)scV.size(), dperm.get(), dpoff.get(), dvs.get(), arena.get(), dV.get());
    // norm_tensor(U) (src/siteTensor_misc.c:244-250)
    DevBuf<double> dnrm(1);
    hipLaunchKernelGGL(sumsq_kernel, dim3(1), dim3(1024), 0, s, dU.get(), ubb[nU], dnrm.get());
    double nrm2 = 0;
    T3_HIP_CHECK(hipMemcpyAsync(&nrm2, dnrm.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    R.norm_U = std::sqrt(nrm2);
    if (ubb[nU] && R.norm_U > 0)
        hipLaunchKernelGGL(scale_kernel, dim3(gridn((ubb[nU] + 255) / 256)), dim3(256), 0, s, dU.get(),
                           ubb[nU], 1.0 / R.norm_U);
    T3_HIP_CHECK(hipEventRecord(e1, s));
    std::vector<double> hu(ubb[nU]), hv(vbb[nV]);
    if (!hu.empty())
        T3_HIP_CHECK(hipMemcpyAsync(hu.data(), dU.get(), hu.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    if (!hv.empty())
        T3_HIP_CHECK(hipMemcpyAsync(hv.data(), dV.get(), hv.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    T3_HIP_CHECK(hipGetLastError());
    if (work) {
        work->flops = flops;
        work->bytes = 8.0 * (double)(A.size() + hu.size() + hv.size());
        work->kernel_ms = ms;
    }

    // remove the sectors without a kept value (adapt_UV_tensors_and_kick_empties)
    std::vector<int> newid(nsec, -1);
    int cnt = 0;
    R.bond = cut;
    R.bond.irreps.clear();
    R.bond.dims.clear();
    R.bond.fcidims.clear();
    R.bond.totaldims = 0;
    for (int ss = 0; ss < nsec; ++ss) {
        if (R.kept[ss] == 0) continue;
        newid[ss] = cnt++;
        R.bond.irreps.push_back(cut.irreps[ss]);
        R.bond.dims.push_back(R.kept[ss]);
        if (!cut.fcidims.empty()) R.bond.fcidims.push_back(cut.fcidims[ss]);
        R.bond.totaldims += R.kept[ss];
    }
    R.bond.nrSecs = cnt;
    auto reform = [&](const std::vector<QN>& qn, int nsites, int site, const Symsecs* ss3, int bond,
                      const std::vector<int64_t>& bb, const std::vector<double>& tel, SiteTensor* out) {
        const int nbl = (int)bb.size() - 1;
        int od[3] = {ss3[0].nrSecs, ss3[1].nrSecs, ss3[2].nrSecs};
        int nd[3] = {od[0], od[1], od[2]};
        nd[bond] = cnt;
        out->nrsites = nsites;
        out->beginblock.assign(1, 0);
        out->qnumbers.clear();
        out->tel.clear();
        for (int b = 0; b < nbl; ++b) {
            const int64_t n = bb[b + 1] - bb[b];
            if (n == 0) continue;
            const QN q = qn[(size_t)b * nsites + site];
            int id[3] = {(int)(q % od[0]), (int)((q / od[0]) % od[1]), (int)(q / ((QN)od[0] * od[1]))};
            id[bond] = newid[id[bond]];
            if (id[bond] < 0) throw std::runtime_error("split: kept block in a removed sector");
            for (int i = 0; i < nsites; ++i) {
                QN v = qn[(size_t)b * nsites + i];
                if (i == site) v = (QN)id[0] + (QN)id[1] * nd[0] + (QN)id[2] * nd[0] * nd[1];
                out->qnumbers.push_back(v);
            }
            out->tel.insert(out->tel.end(), tel.begin() + bb[b], tel.begin() + bb[b + 1]);
            out->beginblock.push_back(out->beginblock.back() + n);
        }
        out->nrblocks = (int)out->beginblock.size() - 1;
    };
    reform(P.U_qn, P.U_nrsites, csiteU, in.ss[P.id_csite], P.id_cbond, ubb, hu, &R.U);
    reform(P.V_qn, 1, 0, sV, P.id_bond, vbb, hv, &R.V);
    return R;
}

}  // namespace t3
