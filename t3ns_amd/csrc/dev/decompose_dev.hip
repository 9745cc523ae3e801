// SVD split on the device (src/siteTensor_decompose.c:587-1444).
#include "decompose_dev.hpp"
#include "tile_gemm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
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

// Small matrices (at most kJSmallCols columns and kJSmallElems elements)
// run every round of every sweep in one workgroup: the pairs of a round
// touch disjoint columns, so each wave rotates one pair at a time (its dot
// products reduced across the wave), a barrier separates the rounds, and
// the workgroup stops after a sweep that rotated nothing (the rotations are
// those of the one-launch-per-round schedule: a matrix that has converged
// rotates nothing in later sweeps).  Larger matrices keep one launch per
// round, where every pair of a round has a workgroup of its own.
constexpr int kJSmallThreads = 512;
constexpr int kJSmallCols = 256;
constexpr int64_t kJSmallElems = 1 << 16;

__global__ void __launch_bounds__(kJSmallThreads)
    jacobi_small_kernel(const JacobiMat* __restrict__ mats, const int* __restrict__ list, int nlist,
                        int max_sweeps, double tol, double* arena, int* sweeps_out) {
    __shared__ int rotated[2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
        const JacobiMat M = mats[list[li]];
        const int np = M.n + (M.n & 1);
        int sweep = 0;
        for (; sweep < max_sweeps; ++sweep) {
            // two flags: the next sweep clears the other one while every
            // thread may still be reading this one
            if (threadIdx.x == 0) rotated[sweep & 1] = 0;
            __syncthreads();
            for (int r = 0; r < np - 1; ++r) {
                for (int k = wave; k < np / 2; k += nw) {
                    int p, q;
                    round_pair(np, r, k, &p, &q);
                    if (p >= M.n || q >= M.n) continue;
                    if (p > q) {
                        const int t = p;
                        p = q;
                        q = t;
                    }
                    double* bp = arena + M.b + (int64_t)p * M.m;
                    double* bq = arena + M.b + (int64_t)q * M.m;
                    // the sums of jacobi_round_kernel in its order: lane l
                    // runs the chains of its threads l + 64 j (rows l + 64 j
                    // + kJThreads s), then the same pairwise tree
                    double a4[4] = {0, 0, 0, 0}, b4[4] = {0, 0, 0, 0}, g4[4] = {0, 0, 0, 0};
                    for (int i0 = 0; i0 < M.m; i0 += kJThreads)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int i = i0 + lane + 64 * j;
                            if (i < M.m) {
                                const double x = bp[i], y = bq[i];
                                a4[j] = fma(x, x, a4[j]);
                                b4[j] = fma(y, y, b4[j]);
                                g4[j] = fma(x, y, g4[j]);
                            }
                        }
                    double a = (a4[0] + a4[2]) + (a4[1] + a4[3]);
                    double b = (b4[0] + b4[2]) + (b4[1] + b4[3]);
                    double g = (g4[0] + g4[2]) + (g4[1] + g4[3]);
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        a += __shfl_down(a, o);
                        b += __shfl_down(b, o);
                        g += __shfl_down(g, o);
                    }
                    a = __shfl(a, 0);
                    b = __shfl(b, 0);
                    g = __shfl(g, 0);
                    // the rotation test and angle of jacobi_round_kernel
                    if (g != 0 && fabs(g) > tol * sqrt((double)M.m) * sqrt(a * b)) {
                        const double zeta = (b - a) / (2 * g);
                        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                        const double c = 1 / sqrt(1 + t * t), s = c * t;
                        for (int i = lane; i < M.m; i += 64) {
                            const double x = bp[i], y = bq[i];
                            bp[i] = c * x - s * y;
                            bq[i] = s * x + c * y;
                        }
                        double* vp = arena + M.v + (int64_t)p * M.n;
                        double* vq = arena + M.v + (int64_t)q * M.n;
                        for (int i = lane; i < M.n; i += 64) {
                            const double x = vp[i], y = vq[i];
                            vp[i] = c * x - s * y;
                            vq[i] = s * x + c * y;
                        }
                        if (lane == 0) rotated[sweep & 1] = 1;
                    }
                }
                __syncthreads();
            }
            if (!rotated[sweep & 1]) break;
        }
        if (threadIdx.x == 0) atomicMax(sweeps_out, sweep + 1);
        __syncthreads();
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
    // T3NS_JACOBI_SMALL_COLS lowers the small path's column limit (0: every
    // matrix takes the per-round path; both give the same bits)
    int small_cols = kJSmallCols;
    if (const char* e = std::getenv("T3NS_JACOBI_SMALL_COLS")) small_cols = std::min(kJSmallCols, std::atoi(e));
    std::vector<JItem> items;
    std::vector<int> small;
    int maxround = 0;
    for (size_t m = 0; m < mats.size(); ++m) {
        if (mats[m].m < mats[m].n) throw std::invalid_argument("jacobi: rows < columns");
        if (mats[m].n <= small_cols && (int64_t)mats[m].m * mats[m].n <= kJSmallElems) {
            small.push_back((int)m);
            continue;
        }
        const int np = mats[m].n + (mats[m].n & 1);
        for (int k = 0; k < np / 2; ++k) items.push_back(JItem{(int)m, k});
        maxround = std::max(maxround, np - 1);
    }
    DevBuf<int> dsmall;
    dsmall.upload(small, s);
    DevBuf<JItem> di;
    di.upload(items, s);
    DevBuf<int> flag(2);  // rotation flag of the per-round schedule, sweeps of the small matrices
    hipLaunchKernelGGL(zero_kernel<int>, dim3(1), dim3(64), 0, s, flag.get(), (int64_t)2);
    hipLaunchKernelGGL(identity_kernel, dim3(gridn((int64_t)mats.size())), dim3(256), 0, s, dm.get(),
                       (int)mats.size(), arena);
    if (!small.empty())
        hipLaunchKernelGGL(jacobi_small_kernel, dim3(gridn((int64_t)small.size())), dim3(kJSmallThreads), 0,
                           s, dm.get(), dsmall.get(), (int)small.size(), max_sweeps, tol, arena,
                           flag.get() + 1);
    int sweep = 0;
    if (!items.empty()) {
        for (; sweep < max_sweeps; ++sweep) {
            hipLaunchKernelGGL(zero_kernel<int>, dim3(1), dim3(64), 0, s, flag.get(), (int64_t)1);
            for (int r = 0; r < maxround; ++r)
                hipLaunchKernelGGL(jacobi_round_kernel, dim3(gridn((int64_t)items.size())), dim3(kJThreads),
                                   0, s, dm.get(), di.get(), (int)items.size(), r, tol, arena, flag.get());
            int h = 0;
            T3_HIP_CHECK(hipMemcpyAsync(&h, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            T3_HIP_CHECK(hipStreamSynchronize(s));
            if (!h) break;
        }
    }
    int sweeps = items.empty() ? 0 : sweep + 1;
    if (!small.empty()) {
        int h = 0;
        T3_HIP_CHECK(hipMemcpyAsync(&h, flag.get() + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        sweeps = std::max(sweeps, h);
    }
    T3_HIP_CHECK(hipGetLastError());
    return sweeps;
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

    // Null directions of transposed sectors: a kept singular value at most
    // 1e-12 of the sector's largest has no direction of its own (its rotated
    // column is rounding noise or zero, and dividing by the value would not
    // give a unit vector).  Such columns are replaced, in singular-value
    // order, by the first unit vector e_c that keeps at least half its norm
    // after two Gram-Schmidt passes against the sector's other good or
    // completed columns (normalised), so V keeps orthonormal rows as
    // dgesdd's VT does.  Rare: only sectors with such values kept are
    // brought to the host.
    std::vector<std::vector<char>> unit(nsec);
    for (int ss = 0; ss < nsec; ++ss) {
        const int mi = mat_of[ss];
        if (mi < 0 || !trans[ss] || R.kept[ss] == 0) continue;
        const std::vector<double>& sg = R.sing[ss];
        const double thr = 1e-12 * (sg.empty() ? 0.0 : sg[0]);
        const int n = mats[mi].n, m = mats[mi].m;
        bool need = false;
        for (int t = 0; t < R.kept[ss]; ++t) need = need || sg[t] <= thr;
        if (!need) continue;
        std::vector<double> B((size_t)m * n);
        T3_HIP_CHECK(hipMemcpyAsync(B.data(), arena.get() + mats[mi].b, B.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<char> ok(n, 0);
        std::vector<std::vector<double>> rows(n);  // unit rows by singular-value rank
        for (int t = 0; t < n; ++t)
            if (sg[t] > thr) {
                ok[t] = 1;
                const double* c = B.data() + (size_t)order[ss][t] * m;
                rows[t].assign(c, c + m);
                for (double& x : rows[t]) x /= sg[t];
            }
        unit[ss].assign(R.kept[ss], 0);
        std::vector<double> v(m);
        for (int t = 0; t < R.kept[ss]; ++t) {
            if (ok[t]) continue;
            for (int c = 0; c < m; ++c) {
                std::fill(v.begin(), v.end(), 0.0);
                v[c] = 1.0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int r = 0; r < n; ++r) {
                        if (!ok[r]) continue;
                        double d = 0;
                        for (int i = 0; i < m; ++i) d += rows[r][i] * v[i];
                        for (int i = 0; i < m; ++i) v[i] -= d * rows[r][i];
                    }
                double nr = 0;
                for (double x : v) nr += x * x;
                nr = std::sqrt(nr);
                if (nr > 0.5) {
                    for (double& x : v) x /= nr;
                    rows[t] = v;
                    ok[t] = 1;
                    unit[ss][t] = 1;
                    std::copy(v.begin(), v.end(), B.begin() + (size_t)order[ss][t] * m);
                    break;
                }
            }
        }
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + mats[mi].b, B.data(), B.size() * sizeof(double),
                                    hipMemcpyHostToDevice, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
    }

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
            const bool u1 = !unit[ss].empty() && unit[ss][t];  // completed: already a unit vector
            vscale.push_back(trans[ss] ? (u1 ? 1.0 : (sg > 0 ? 1.0 / sg : 0.0)) : 1.0);
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

// ---------------------------------------------------------------------------
// qr_step: Householder QR per sector (dgeqrf + dorgqr semantics) and R
// multiplied into the centre.
// ---------------------------------------------------------------------------
namespace t3 {

namespace {

struct QRMat {
    int64_t a, tau, q, r;  // offsets: matrix (m x n), tau (k), Q (m x k), R (k x n)
    int m, n, k, pad;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One workgroup per matrix: Householder vectors below the diagonal, beta on
// it (LAPACK dlarfg convention: beta = -sign(alpha) * norm), then R (upper
// triangle) and the explicit Q of the first k columns.
__global__ void __launch_bounds__(256)
    householder_qr_kernel(const QRMat* __restrict__ mats, int nm, double* arena) {
    __shared__ double red[256];
    __shared__ double sh_tau, sh_scale;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int mi = blockIdx.x; mi < nm; mi += gridDim.x) {
        const QRMat Q = mats[mi];
        double* A = arena + Q.a;
        double* tau = arena + Q.tau;
        const int m = Q.m, n = Q.n;
        for (int k = 0; k < Q.k; ++k) {
            double* col = A + (int64_t)k * m;
            double x2 = 0;
            for (int i = k + 1 + threadIdx.x; i < m; i += 256) x2 = fma(col[i], col[i], x2);
            red[threadIdx.x] = x2;
            __syncthreads();
            for (int st = 128; st > 0; st >>= 1) {
                if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                const double alpha = col[k], xn2 = red[0];
                if (xn2 == 0) {
                    sh_tau = 0;
                    sh_scale = 0;
                } else {
                    const double nrm = sqrt(alpha * alpha + xn2);
                    const double beta = alpha >= 0 ? -nrm : nrm;
                    sh_tau = (beta - alpha) / beta;
                    sh_scale = 1 / (alpha - beta);
                    col[k] = beta;
                }
                tau[k] = sh_tau;
            }
            __syncthreads();
            const double t = sh_tau, sc = sh_scale;
            if (t != 0)
                for (int i = k + 1 + threadIdx.x; i < m; i += 256) col[i] *= sc;
            __syncthreads();
            // apply H = I - tau v v^T (v_k = 1) to columns k+1..n-1, one wave per column
            if (t != 0)
                for (int j = k + 1 + wave; j < n; j += 4) {
                    double* cj = A + (int64_t)j * m;
                    double w = lane == 0 ? cj[k] : 0.0;
                    for (int i = k + 1 + lane; i < m; i += 64) w = fma(col[i], cj[i], w);
                    w = wave_sum(w);
                    if (lane == 0) cj[k] -= t * w;
                    for (int i = k + 1 + lane; i < m; i += 64) cj[i] -= t * col[i] * w;
                }
            __syncthreads();
        }
        // R: the upper triangle of the first k rows (copy_to_R)
        double* R = arena + Q.r;
        for (int64_t e = threadIdx.x; e < (int64_t)Q.k * n; e += 256) {
            const int i = (int)(e % Q.k), j = (int)(e / Q.k);
            R[e] = i <= j ? A[i + (int64_t)j * m] : 0.0;
        }
        // Q = H_0 ... H_{k-1} applied to the first k columns of the identity
        double* q = arena + Q.q;
        for (int64_t e = threadIdx.x; e < (int64_t)m * Q.k; e += 256) {
            const int i = (int)(e % m), j = (int)(e / m);
            q[e] = i == j ? 1.0 : 0.0;
        }
        __syncthreads();
        for (int kk = Q.k - 1; kk >= 0; --kk) {
            const double t = tau[kk];
            if (t == 0) continue;  // uniform
            const double* v = A + (int64_t)kk * m;
            for (int j = kk + wave; j < Q.k; j += 4) {
                double* cj = q + (int64_t)j * m;
                double w = lane == 0 ? cj[kk] : 0.0;
                for (int i = kk + 1 + lane; i < m; i += 64) w = fma(v[i], cj[i], w);
                w = wave_sum(w);
                if (lane == 0) cj[kk] -= t * w;
                for (int i = kk + 1 + lane; i < m; i += 64) cj[i] -= t * v[i] * w;
            }
            __syncthreads();
        }
    }
}

// A block (before, N, after) into rows row0 + b + a * before, columns 0..N-1
// of its sector matrix, and back from Q (before, K, after).
struct QRMove {
    int64_t blk, mat, row0;
    int64_t before, after;
    int64_t ld, cols;  // matrix rows (leading dimension), columns moved
};

__global__ void qr_gather_kernel(const QRMove* __restrict__ mv, int n, const double* __restrict__ src,
                                 double* arena) {
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const QRMove M = mv[b];
        const int64_t tot = M.before * M.cols * M.after;
        for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
            const int64_t x = e % M.before, c = (e / M.before) % M.cols, y = e / (M.before * M.cols);
            arena[M.mat + M.row0 + x + y * M.before + c * M.ld] = src[M.blk + e];
        }
    }
}

__global__ void qr_scatter_kernel(const QRMove* __restrict__ mv, int n, const double* arena,
                                  double* __restrict__ dst) {
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const QRMove M = mv[b];
        const int64_t tot = M.before * M.cols * M.after;
        for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
            const int64_t x = e % M.before, c = (e / M.before) % M.cols, y = e / (M.before * M.cols);
            dst[M.blk + e] = arena[M.mat + M.row0 + x + y * M.before + c * M.ld];
        }
    }
}

void mode_tiles_qr(std::vector<tile::GemmTile>& out, int64_t x, const int64_t n[3], int leg, int64_t o,
                   int e, int64_t y) {
    tile::GemmTile g;
    auto push = [&]() {
        for (int a = 0; a < tile::tiles_of(g.M); ++a)
            for (int b = 0; b < tile::tiles_of(g.N); ++b) {
                g.ti = a;
                g.tj = b;
                out.push_back(g);
            }
    };
    if (leg == 0) {
        g.p = tile::Strided{o, 1, e};
        g.q = tile::Strided{x, 1, (int)n[0]};
        g.M = e;
        g.N = (int)(n[1] * n[2]);
        g.K = (int)n[0];
        g.x_off = y;
        g.ldx = e;
        push();
    } else if (leg == 2) {
        g.p = tile::Strided{x, 1, (int)(n[0] * n[1])};
        g.q = tile::Strided{o, e, 1};
        g.M = (int)(n[0] * n[1]);
        g.N = e;
        g.K = (int)n[2];
        g.x_off = y;
        g.ldx = g.M;
        push();
    } else {
        for (int64_t l = 0; l < n[2]; ++l) {
            g.p = tile::Strided{x + l * n[0] * n[1], 1, (int)n[0]};
            g.q = tile::Strided{o, e, 1};
            g.M = (int)n[0];
            g.N = e;
            g.K = (int)n[1];
            g.x_off = y + l * n[0] * e;
            g.ldx = (int)n[0];
            push();
        }
    }
}

}  // namespace

QRResult run_qr(const QRInput& in, const QRPlan& P, hipStream_t s, RenormWork* work) {
    const SiteTensor& A = *in.A;
    const SiteTensor& C = *in.C;
    const int nsec = (int)P.sec.size();
    QRResult R;
    R.dims.assign(nsec, 0);
    // arena: [A | C | per sector: matrix, tau, Q, R | B]
    const int64_t a_base = 0, c_base = A.size();
    int64_t top = c_base + C.size();
    std::vector<QRMat> mats;
    std::vector<int> mat_of(nsec, -1);
    double flops = 0;
    for (int ss = 0; ss < nsec; ++ss) {
        const QRSector& S = P.sec[ss];
        R.dims[ss] = (int)S.K;
        if (S.K == 0) continue;
        QRMat q;
        q.m = (int)S.M;
        q.n = (int)S.N;
        q.k = (int)S.K;
        q.pad = 0;
        q.a = top;
        top += S.M * S.N;
        q.tau = top;
        top += S.K;
        q.q = top;
        top += S.M * S.K;
        q.r = top;
        top += S.K * S.N;
        mat_of[ss] = (int)mats.size();
        mats.push_back(q);
        flops += 4.0 * S.M * S.N * S.K;  // factorisation and explicit Q
    }
    // Q blocks: (before, K, after); B blocks: the centre with N -> K
    std::vector<QRMove> gat, sca;
    R.Q.nrsites = 1;
    R.Q.sites[0] = A.sites[0];
    R.Q.nrblocks = A.nrblocks;
    R.Q.qnumbers = A.qnumbers;
    R.Q.beginblock.assign(A.nrblocks + 1, 0);
    for (int b = 0; b < A.nrblocks; ++b) {
        const QRPlan::ABlock& ab = P.ablocks[b];
        const QRSector& S = P.sec[ab.sector];
        const int64_t sz = A.block_size(b) == 0 ? 0 : ab.before * S.K * ab.after;
        R.Q.beginblock[b + 1] = R.Q.beginblock[b] + sz;
    }
    for (int ss = 0; ss < nsec; ++ss) {
        const QRSector& S = P.sec[ss];
        if (mat_of[ss] < 0) continue;
        const QRMat& q = mats[mat_of[ss]];
        for (size_t i = 0; i < S.blocks.size(); ++i) {
            const int b = S.blocks[i];
            const QRPlan::ABlock& ab = P.ablocks[b];
            gat.push_back(QRMove{a_base + A.beginblock[b], q.a, S.rowstart[i], ab.before, ab.after, S.M, S.N});
            sca.push_back(QRMove{R.Q.beginblock[b], q.q, S.rowstart[i], ab.before, ab.after, S.M, S.K});
        }
    }
    // B = C x_bond R
    const int bc = P.bondC;
    R.B.nrsites = 1;
    R.B.sites[0] = C.sites[0];
    R.B.nrblocks = C.nrblocks;
    R.B.qnumbers = C.qnumbers;
    R.B.beginblock.assign(C.nrblocks + 1, 0);
    std::vector<tile::GemmTile> tiles;
    const int64_t b_base = top;
    for (int b = 0; b < C.nrblocks; ++b) {
        int id[3];
        const QN n0 = in.ssC[0].nrSecs, n1 = in.ssC[1].nrSecs;
        id[0] = (int)(C.qnumbers[b] % n0);
        id[1] = (int)((C.qnumbers[b] / n0) % n1);
        id[2] = (int)(C.qnumbers[b] / (n0 * n1));
        const int ss = id[bc];
        const int64_t K = (ss < nsec) ? P.sec[ss].K : 0;
        int64_t nd[3];
        for (int i = 0; i < 3; ++i) nd[i] = in.ssC[i].dims[id[i]];
        const int64_t sz = (C.block_size(b) == 0 || K == 0) ? 0 : C.block_size(b) / nd[bc] * K;
        R.B.beginblock[b + 1] = R.B.beginblock[b] + sz;
        if (sz == 0) continue;
        if (nd[bc] != P.sec[ss].N) throw std::runtime_error("qr: centre and tensor disagree on the bond");
        mode_tiles_qr(tiles, c_base + C.beginblock[b], nd, bc, mats[mat_of[ss]].r, (int)K,
                      b_base + R.B.beginblock[b]);
        flops += 2.0 * (double)C.block_size(b) * K;
    }
    top = b_base + R.B.beginblock[C.nrblocks];
    DevBuf<double> arena(std::max<int64_t>(1, top));
    if (A.size())
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + a_base, A.tel.data(), A.size() * sizeof(double),
                                    hipMemcpyHostToDevice, s));
    if (C.size())
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + c_base, C.tel.data(), C.size() * sizeof(double),
                                    hipMemcpyHostToDevice, s));
    DevBuf<QRMat> dm;
    DevBuf<QRMove> dg, dsc;
    DevBuf<tile::GemmTile> dt;
    dm.upload(mats, s);
    dg.upload(gat, s);
    dsc.upload(sca, s);
    dt.upload(tiles, s);
    DevBuf<double> dQ(std::max<int64_t>(1, R.Q.size()));
    // sector matrices start zeroed
    for (const QRMat& q : mats)
        hipLaunchKernelGGL(zero_kernel<double>, dim3(gridn(((int64_t)q.m * q.n + 255) / 256)), dim3(256), 0,
                           s, arena.get() + q.a, (int64_t)q.m * q.n);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    if (!gat.empty())
        hipLaunchKernelGGL(qr_gather_kernel, dim3(gridn((int64_t)gat.size())), dim3(256), 0, s, dg.get(),
                           (int)gat.size(), arena.get(), arena.get());
    if (!mats.empty())
        hipLaunchKernelGGL(householder_qr_kernel, dim3(gridn((int64_t)mats.size())), dim3(256), 0, s,
                           dm.get(), (int)mats.size(), arena.get());
    if (!sca.empty())
        hipLaunchKernelGGL(qr_scatter_kernel, dim3(gridn((int64_t)sca.size())), dim3(256), 0, s, dsc.get(),
                           (int)sca.size(), arena.get(), dQ.get());
    if (!tiles.empty())
        hipLaunchKernelGGL(tile::gemm_tiles_kernel, dim3(gridn((int64_t)tiles.size())), dim3(256), 0, s,
                           dt.get(), (int)tiles.size(), arena.get());
    T3_HIP_CHECK(hipEventRecord(e1, s));
    R.Q.tel.assign(R.Q.size(), 0.0);
    R.B.tel.assign(R.B.size(), 0.0);
    if (R.Q.size())
        T3_HIP_CHECK(hipMemcpyAsync(R.Q.tel.data(), dQ.get(), R.Q.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    if (R.B.size())
        T3_HIP_CHECK(hipMemcpyAsync(R.B.tel.data(), arena.get() + b_base, R.B.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    // singular values of R (R_svd, 550-574): Jacobi on R^T (n x k, n >= k)
    std::vector<JacobiMat> jm;
    int64_t jtop = 0;
    for (int ss = 0; ss < nsec; ++ss) {
        if (mat_of[ss] < 0) continue;
        const QRMat& q = mats[mat_of[ss]];
        JacobiMat J;
        J.m = q.n;
        J.n = q.k;
        J.b = jtop;
        jtop += (int64_t)J.m * J.n;
        J.v = jtop;
        jtop += (int64_t)J.n * J.n;
        jm.push_back(J);
    }
    DevBuf<double> jar(std::max<int64_t>(1, jtop));
    // the R matrices are small (k x n per sector): transpose them on the host
    std::vector<double> hj(std::max<int64_t>(1, jtop), 0.0);
    {
        size_t t = 0;
        std::vector<double> hr;
        for (int ss = 0; ss < nsec; ++ss) {
            if (mat_of[ss] < 0) continue;
            const QRMat& q = mats[mat_of[ss]];
            const JacobiMat& J = jm[t++];
            hr.resize((size_t)q.k * q.n);
            T3_HIP_CHECK(hipMemcpyAsync(hr.data(), arena.get() + q.r, hr.size() * sizeof(double),
                                        hipMemcpyDeviceToHost, s));
            T3_HIP_CHECK(hipStreamSynchronize(s));
            for (int i = 0; i < q.k; ++i)
                for (int j = 0; j < q.n; ++j) hj[J.b + j + (int64_t)i * q.n] = hr[i + (int64_t)j * q.k];
        }
    }
    if (jtop) T3_HIP_CHECK(hipMemcpyAsync(jar.get(), hj.data(), jtop * sizeof(double), hipMemcpyHostToDevice, s));
    jacobi_svd(jar.get(), jm, s, 40, 1e-15);
    std::vector<double> hs(std::max<int64_t>(1, jtop));
    if (jtop) T3_HIP_CHECK(hipMemcpyAsync(hs.data(), jar.get(), jtop * sizeof(double), hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    T3_HIP_CHECK(hipGetLastError());
    R.sing.assign(nsec, {});
    {
        size_t t = 0;
        for (int ss = 0; ss < nsec; ++ss) {
            if (mat_of[ss] < 0) continue;
            const JacobiMat& J = jm[t++];
            std::vector<double> sv;
            for (int j = 0; j < J.n; ++j) {
                double a = 0;
                for (int i = 0; i < J.m; ++i) a += hs[J.b + i + (int64_t)j * J.m] * hs[J.b + i + (int64_t)j * J.m];
                sv.push_back(std::sqrt(a));
            }
            std::sort(sv.begin(), sv.end(), [](double x, double y) { return x > y; });
            R.sing[ss] = sv;
        }
    }
    // calculateCostFunction(S, VonNeumannEntropy, 'A') over the bond's sectors
    R.entropy = 0;
    for (int ss = 0; ss < nsec; ++ss) {
        const int mult = multiplicity(in.sym, in.ssA[P.bondA].irreps[ss].data());
        const double mf = 1 / std::sqrt((double)mult);
        double e = 0;
        for (double v : R.sing[ss]) {
            const double w = (v * mf) * (v * mf);
            e += w < 1e-20 ? 0 : -w * std::log(w);
        }
        R.entropy += mult * e;
    }
    if (work) {
        work->flops = flops;
        work->bytes = 8.0 * (double)(A.size() + C.size() + R.Q.size() + R.B.size());
        work->kernel_ms = ms;
    }
    return R;
}

}  // namespace t3
