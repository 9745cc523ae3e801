// Device kernels of the renormalisation plans: grouped scaled sums of
// operator blocks, empty-block detection and compaction.  All are
// bandwidth-bound (each element is read once per contributing term and
// written once).
#include "renorm_dev.hpp"
#include "../host/host_threads.hpp"
#include "tile_gemm.hpp"

#include <algorithm>
#include <array>
#include <unordered_map>

namespace t3 {

namespace {
constexpr int kThreads = 256;

struct DevGroup {
    int64_t dst, n;
    int task_begin, task_end;
    int flat;  // flat (operator, block) index for the keep flags
    int pad;
};

// out[dst + e] = sum_t pref_t * (factor_t * old[src_t + e]); keep[flat] = any
// element with |value| >= 1e-20 (the reference's zero test).
__global__ void __launch_bounds__(kThreads)
    append_kernel(const DevGroup* __restrict__ groups, int ngroups,
                  const int64_t* __restrict__ src, const double* __restrict__ factor,
                  const double* __restrict__ pref, const double* __restrict__ old,
                  double* __restrict__ out, unsigned char* __restrict__ keep) {
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const DevGroup g = groups[gi];
        int any = 0;
        for (int64_t e = threadIdx.x; e < g.n; e += kThreads) {
            double v = 0;
            for (int t = g.task_begin; t < g.task_end; ++t) v += pref[t] * (factor[t] * old[src[t] + e]);
            out[g.dst + e] = v;
            any |= fabs(v) >= 1e-20;
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) keep[g.flat] = (unsigned char)(any != 0);
    }
}

__global__ void __launch_bounds__(kThreads)
    gather_blocks_kernel(const int64_t* __restrict__ src, const int64_t* __restrict__ dst,
                         const int64_t* __restrict__ n, int nb, const double* __restrict__ in,
                         double* __restrict__ out) {
    for (int b = blockIdx.x; b < nb; b += gridDim.x)
        for (int64_t e = threadIdx.x; e < n[b]; e += kThreads) out[dst[b] + e] = in[src[b] + e];
}

template <typename T>
__global__ void fill_zero(T* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

int grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, n)); }
}  // namespace

// dst groups = sum_t pref_t * (factor_t * src[src_t]) over a device source,
// then the blocks left all zero are dropped (src/sparseblocks.c:73-107) and
// the result copied to the host.
static ROperators finish_sum(const ROperators& nr, const std::vector<AppendGroup>& groups,
                             const std::vector<AppendTask>& tasks, const double* d_src, hipStream_t s,
                             float* kernel_ms = nullptr, DevArena* keep_dev = nullptr) {
    const int64_t nnew = nr.total_size();
    // flat block index of every (op, block)
    std::vector<int> flat_of_op(nr.nrops + 1, 0);
    for (int k = 0; k < nr.nrops; ++k) flat_of_op[k + 1] = flat_of_op[k] + nr.nblocks_op(k);
    const int nflat = flat_of_op[nr.nrops];
    std::vector<DevGroup> hg(groups.size());
    for (size_t i = 0; i < groups.size(); ++i) {
        const AppendGroup& g = groups[i];
        hg[i] = DevGroup{g.dst, g.n, g.task_begin, g.task_end, flat_of_op[g.op] + g.block, 0};
    }
    std::vector<int64_t> tsrc(tasks.size());
    std::vector<double> tfac(tasks.size()), tpref(tasks.size());
    for (size_t i = 0; i < tasks.size(); ++i) {
        tsrc[i] = tasks[i].src;
        tfac[i] = tasks[i].factor;
        tpref[i] = tasks[i].pref;
    }
    DevBuf<DevGroup> dg;
    DevBuf<int64_t> dsrc;
    DevBuf<double> dfac, dpref, dnew;
    DevBuf<unsigned char> dkeep;
    dg.upload(hg, s);
    dsrc.upload(tsrc, s);
    dfac.upload(tfac, s);
    dpref.upload(tpref, s);
    dnew.alloc(std::max<int64_t>(1, nnew));
    dkeep.alloc(std::max(1, nflat));
    hipLaunchKernelGGL(fill_zero<unsigned char>, dim3(grid((nflat + 255) / 256)), dim3(256), 0, s,
                       dkeep.get(), (int64_t)nflat);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    if (!hg.empty())
        hipLaunchKernelGGL(append_kernel, dim3(grid((int64_t)hg.size())), dim3(kThreads), 0, s,
                           dg.get(), (int)hg.size(), dsrc.get(), dfac.get(), dpref.get(), d_src,
                           dnew.get(), dkeep.get());
    T3_HIP_CHECK(hipEventRecord(e1, s));
    std::vector<unsigned char> keep(nflat);
    if (nflat)
        T3_HIP_CHECK(hipMemcpyAsync(keep.data(), dkeep.get(), nflat, hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    const Compaction cpt = compact_blocks(nr, keep);
    ROperators out = nr;
    out.beginblock = cpt.beginblock;
    out.op_offset = cpt.op_offset;
    const int64_t ncomp = cpt.op_offset[nr.nrops];
    out.tel.assign(ncomp, 0.0);
    if (ncomp) {
        DevBuf<int64_t> cs, cd, cn;
        auto dcomp = std::make_shared<DevBuf<double>>(ncomp);
        cs.upload(cpt.src, s);
        cd.upload(cpt.dst, s);
        cn.upload(cpt.n, s);
        hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid((int64_t)cpt.n.size())), dim3(kThreads), 0,
                           s, cs.get(), cd.get(), cn.get(), (int)cpt.n.size(), dnew.get(),
                           dcomp->get());
        T3_HIP_CHECK(hipMemcpyAsync(out.tel.data(), dcomp->get(), ncomp * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        if (keep_dev) *keep_dev = std::move(dcomp);
    }
    T3_HIP_CHECK(hipGetLastError());
    return out;
}

ROperators run_append(const AppendPlan& plan, const ROperators& old, hipStream_t s, RenormWork* work,
                      const double* dev_old, DevArena* keep) {
    DevBuf<double> dold;
    if (!dev_old) {
        dold.upload(old.tel, s);
        if (old.tel.empty()) dold.alloc(1);
    }
    float ms = 0;
    ROperators out = finish_sum(plan.nr, plan.groups, plan.tasks, dev_old ? dev_old : dold.get(), s,
                                &ms, keep);
    if (work) {
        // each task reads its source block once; every new block is written once
        double flops = 0, rd = 0;
        for (const AppendGroup& g : plan.groups) {
            rd += (double)g.n * (g.task_end - g.task_begin);
            flops += 3.0 * g.n * (g.task_end - g.task_begin);
        }
        work->flops = flops;
        work->bytes = 8.0 * (rd + (double)plan.nr.total_size());
        work->kernel_ms = ms;
    }
    return out;
}

}  // namespace t3

// ---------------------------------------------------------------------------
// update_rOperators_physical: U[dst] = sum_t pref_t * A_t O_t B_t.
// ---------------------------------------------------------------------------
namespace t3 {

namespace {

using tile::GemmTile;
using tile::kTile;
using tile::Strided;
using tile::tiles_of;

// Phase 2: U tile = sum_t pref_t * R_t S_t.
struct SumTile {
    int64_t u_off;
    int M, N;
    int ti, tj;
    int task_begin, task_end;
    int flat;
    int pad;
};
struct SumTask {
    Strided r, s;  // s as (k, j)
    int K;
    int pad;
    double pref;
};

// Output blocks of at most kDotEdge x kDotEdge states (a 16 x 16 strip
// would be mostly padding) take their own kernel: the whole workgroup
// splits the (contribution, k) reduction over its threads, neighbouring
// threads taking neighbouring k, and sums the per-thread partials in a
// fixed order (within each wave, then over the waves).
constexpr int kDotEdge = 8;
__host__ __device__ inline bool dot_tile(int M, int N) { return M <= kDotEdge && N <= kDotEdge; }

__device__ __forceinline__ double wave_total(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(256)
    pupdate_dot_kernel(const SumTile* __restrict__ tiles, int ntiles,
                       const SumTask* __restrict__ tasks, const double* __restrict__ base,
                       double* __restrict__ out, unsigned int* __restrict__ keep) {
    __shared__ double red[4][kDotEdge * kDotEdge];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const SumTile g = tiles[it];
        double acc[kDotEdge][kDotEdge] = {};
        for (int t = g.task_begin; t < g.task_end; ++t) {
            const SumTask tk = tasks[t];
            int64_t ro[kDotEdge], so[kDotEdge];  // rows / columns past the block repeat the last
#pragma unroll
            for (int a = 0; a < kDotEdge; ++a) {
                ro[a] = tk.r.off + (int64_t)min(a, g.M - 1) * tk.r.si;
                so[a] = tk.s.off + (int64_t)min(a, g.N - 1) * tk.s.sk;
            }
            for (int k = threadIdx.x; k < tk.K; k += 256) {
                double r[kDotEdge], q[kDotEdge];
#pragma unroll
                for (int a = 0; a < kDotEdge; ++a) {
                    r[a] = base[ro[a] + (int64_t)k * tk.r.sk] * tk.pref;
                    q[a] = base[so[a] + (int64_t)k * tk.s.si];
                }
#pragma unroll
                for (int a = 0; a < kDotEdge; ++a)
#pragma unroll
                    for (int b = 0; b < kDotEdge; ++b) acc[a][b] = fma(r[a], q[b], acc[a][b]);
            }
        }
#pragma unroll
        for (int a = 0; a < kDotEdge; ++a)
#pragma unroll
            for (int b = 0; b < kDotEdge; ++b) {
                const double v = wave_total(acc[a][b]);
                if (lane == 0) red[w][a * kDotEdge + b] = v;
            }
        __syncthreads();
        int any = 0;
        if (threadIdx.x < kDotEdge * kDotEdge) {
            const int a = threadIdx.x / kDotEdge, b = threadIdx.x % kDotEdge;
            const double v = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                             red[3][threadIdx.x];
            if (a < g.M && b < g.N) {
                out[g.u_off + a + (int64_t)b * g.M] = v;
                any = fabs(v) >= 1e-20;
            }
        }
        any = __syncthreads_or(any);  // also keeps red until every thread has read it
        if (threadIdx.x == 0 && any) atomicOr(&keep[g.flat], 1u);
    }
}

__global__ void __launch_bounds__(256)
    pupdate_sum_kernel(const SumTile* __restrict__ tiles, int ntiles,
                       const SumTask* __restrict__ tasks, const double* __restrict__ base,
                       double* __restrict__ out, unsigned int* __restrict__ keep) {
    __shared__ double Ps[tile::kKStep][kTile + 1];
    __shared__ double Qs[tile::kKStep][kTile + 1];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // one wave's strips of tile g (rows i0 .. of the tile), the prefactor
    // scaling the left operand of each contribution; returns "any kept"
    auto strips = [&](const SumTile& g, int i_first, int i_step) {
        const int ti0 = g.ti * kTile, j0 = g.tj * kTile, nq = tile::tile_nq(g.N, j0);
        int any = 0;
        for (int i0 = ti0 + i_first; i0 < min(g.M, ti0 + kTile); i0 += i_step) {
            tile::v4d u[4] = {};
            for (int t = g.task_begin; t < g.task_end; ++t) {
                const SumTask tk = tasks[t];
                tile::wave_strip(base, tk.r, tk.s, g.M, g.N, tk.K, i0, j0, nq, tk.pref, u);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = j0 + 16 * q + (lane & 15);
                if (q >= nq || j >= g.N) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + (lane >> 4) + 4 * v;
                    if (i < g.M) {
                        out[g.u_off + i + (int64_t)j * g.M] = u[q][v];
                        any |= fabs(u[q][v]) >= 1e-20;
                    }
                }
            }
        }
        return any;
    };
    for (int it = blockIdx.x * 4; it < ntiles; it += gridDim.x * 4) {
        // tiles taken four at a time: when all four take the per-wave
        // strips, each wave computes one tile, else the workgroup computes
        // them one after another
        int mine = 0;
        if (it + w < ntiles) {
            const SumTile g = tiles[it + w];
            mine = g.task_end > g.task_begin &&
                   tile::tile_staged(g.M, g.N, tasks[g.task_begin].K, g.ti * kTile, g.tj * kTile);
        }
        if (!__syncthreads_or(mine)) {
            if (it + w < ntiles) {
                const SumTile g = tiles[it + w];
                const int any = __any(strips(g, 0, 16));
                if (lane == 0 && any) atomicOr(&keep[g.flat], 1u);
            }
            continue;
        }
        for (int ti = it; ti < min(it + 4, ntiles); ++ti) {
            const SumTile g = tiles[ti];
            const int ti0 = g.ti * kTile, j0 = g.tj * kTile;
            int any = 0;
            if (g.task_end > g.task_begin &&
                tile::tile_staged(g.M, g.N, tasks[g.task_begin].K, ti0, j0)) {
                // full tiles: LDS slabs, vector FMAs
                double u[4][4] = {};
                for (int t = g.task_begin; t < g.task_end; ++t) {
                    const SumTask tk = tasks[t];
                    double acc[4][4] = {};
                    for (int k0 = 0; k0 < tk.K; k0 += tile::kKStep)
                        tile::tile_slab(base, tk.r, tk.s, g.M, g.N, tk.K, ti0, j0, k0, Ps, Qs, acc);
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) u[a][b] = fma(tk.pref, acc[a][b], u[a][b]);
                }
                const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int j = j0 + ty + 16 * b;
                    if (j >= g.N) continue;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int i = ti0 + tx + 16 * a;
                        if (i < g.M) {
                            out[g.u_off + i + (int64_t)j * g.M] = u[a][b];
                            any |= fabs(u[a][b]) >= 1e-20;
                        }
                    }
                }
            } else {
                any = strips(g, 16 * w, 64);  // the four waves' strips of the tile
            }
            any = __syncthreads_or(any);
            if (threadIdx.x == 0 && any) atomicOr(&keep[g.flat], 1u);
        }
    }
}

// Small-output tiles first (their count returned), then the others.
int partition_sums(std::vector<SumTile>& sums) {
    auto mid = std::stable_partition(sums.begin(), sums.end(),
                                     [](const SumTile& g) { return dot_tile(g.M, g.N); });
    return (int)(mid - sums.begin());
}

void launch_sums(const SumTile* d, int ndot, int n, const SumTask* tasks, const double* base,
                 double* out, unsigned int* keep, hipStream_t s) {
    if (ndot)
        hipLaunchKernelGGL(pupdate_dot_kernel, dim3(grid(ndot)), dim3(256), 0, s, d, ndot, tasks, base,
                           out, keep);
    if (n > ndot)
        hipLaunchKernelGGL(pupdate_sum_kernel, dim3(grid((n - ndot + 3) / 4)), dim3(256), 0, s, d + ndot,
                           n - ndot, tasks, base, out, keep);
}

}  // namespace

ROperators run_pupdate(const PUpdatePlan& plan, const ROperators& pops, const SiteTensor& T,
                       hipStream_t s, RenormWork* work, const double* dev_pops, DevArena* keep_dev) {
    const ROperators& nr = plan.nr;
    const int64_t nT = T.size(), nold = pops.total_size(), nnew = nr.total_size();
    std::vector<int> flat_of_op(nr.nrops + 1, 0);
    for (int k = 0; k < nr.nrops; ++k) flat_of_op[k + 1] = flat_of_op[k] + nr.nblocks_op(k);
    const int nflat = flat_of_op[nr.nrops];

    // Device arena: [T | old operators | workspace of the first products].
    const int64_t oldoff = nT, workoff = nT + nold;
    std::vector<GemmTile> first;
    std::vector<SumTile> sums;
    std::vector<SumTask> stasks;
    int64_t wsize = 0;
    double flops = 0;
    for (const PUpdateGroup& g : plan.groups) {
        SumTile st;
        st.u_off = g.dst;
        st.task_begin = (int)stasks.size();
        st.flat = flat_of_op[g.op] + g.block;
        st.pad = 0;
        int N0 = -1, N1 = -1;
        for (int ti = g.task_begin; ti < g.task_end; ++ti) {
            const PUpdateTask& t = plan.tasks[ti];
            N0 = t.N0;
            N1 = t.N1;
            const int64_t M0 = t.M0, M1 = t.M1;
            // A = Th^T (left) or Th (right), N0 x M0; B = T (left) or T^T
            // (right), M1 x N1; O is M0 x M1 column-major.
            const Strided A = nr.is_left ? Strided{t.th, t.M0, 1} : Strided{t.th, 1, t.N0};
            const Strided B = nr.is_left ? Strided{t.t, 1, t.M1} : Strided{t.t, t.N1, 1};
            const Strided O = Strided{oldoff + t.src, 1, t.M0};
            const double costA = (double)N0 * M0 * M1 + (double)N0 * M1 * N1;
            const double costB = (double)M0 * M1 * N1 + (double)N0 * M0 * N1;
            SumTask sk;
            sk.pref = t.pref;
            sk.pad = 0;
            GemmTile ft;
            ft.x_off = workoff + wsize;
            if (costA <= costB) {
                // W = A O (N0 x M1), then W B
                ft.p = A;
                ft.q = O;
                ft.ldx = N0;
                ft.M = N0;
                ft.N = (int)M1;
                ft.K = (int)M0;
                sk.r = Strided{ft.x_off, 1, N0};
                sk.s = B;
                sk.K = (int)M1;
                wsize += (int64_t)N0 * M1;
                flops += 2 * costA;
            } else {
                // V = O B (M0 x N1), then A V
                ft.p = O;
                ft.q = B;
                ft.ldx = (int)M0;
                ft.M = (int)M0;
                ft.N = N1;
                ft.K = (int)M1;
                sk.r = A;
                sk.s = Strided{ft.x_off, 1, (int)M0};
                sk.K = (int)M0;
                wsize += M0 * N1;
                flops += 2 * costB;
            }
            for (int a = 0; a < tiles_of(ft.M); ++a)
                for (int b = 0; b < tiles_of(ft.N); ++b) {
                    ft.ti = a;
                    ft.tj = b;
                    first.push_back(ft);
                }
            stasks.push_back(sk);
        }
        st.task_end = (int)stasks.size();
        st.M = N0;
        st.N = N1;
        for (int a = 0; a < tiles_of(N0); ++a)
            for (int b = 0; b < tiles_of(N1); ++b) {
                st.ti = a;
                st.tj = b;
                sums.push_back(st);
            }
    }
    if (workoff + wsize > ((int64_t)1 << 33))
        throw std::runtime_error("pUpdate: device workspace above 64 GiB");

    DevBuf<double> arena(std::max<int64_t>(1, workoff + wsize));
    if (nT) T3_HIP_CHECK(hipMemcpyAsync(arena.get(), T.tel.data(), nT * sizeof(double),
                                        hipMemcpyHostToDevice, s));
    if (nold)
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + oldoff, dev_pops ? dev_pops : pops.tel.data(),
                                    nold * sizeof(double),
                                    dev_pops ? hipMemcpyDeviceToDevice :
                                    hipMemcpyHostToDevice, s));
    DevBuf<GemmTile> dfirst;
    DevBuf<SumTile> dsums;
    DevBuf<SumTask> dtasks;
    DevBuf<double> dnew(std::max<int64_t>(1, nnew));
    DevBuf<unsigned int> dkeep(std::max(1, nflat));
    dfirst.upload(first, s);
    const int ndot = partition_sums(sums);
    dsums.upload(sums, s);
    dtasks.upload(stasks, s);
    hipLaunchKernelGGL(fill_zero<unsigned int>, dim3(grid((nflat + 255) / 256)), dim3(256), 0, s,
                       dkeep.get(), (int64_t)nflat);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    if (!first.empty())
        hipLaunchKernelGGL(tile::gemm_tiles_kernel, dim3(grid((int64_t)first.size())), dim3(256), 0,
                           s, dfirst.get(), (int)first.size(), arena.get());
    launch_sums(dsums.get(), ndot, (int)sums.size(), dtasks.get(), arena.get(), dnew.get(),
                dkeep.get(), s);
    T3_HIP_CHECK(hipEventRecord(e1, s));
    std::vector<unsigned int> keep32(nflat);
    if (nflat)
        T3_HIP_CHECK(hipMemcpyAsync(keep32.data(), dkeep.get(), nflat * sizeof(unsigned int),
                                    hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    T3_HIP_CHECK(hipGetLastError());
    if (work) {
        work->flops = flops;
        work->bytes = 8.0 * (double)(nT + nold + nnew);
        work->kernel_ms = ms;
    }

    std::vector<unsigned char> keep(nflat);
    for (int i = 0; i < nflat; ++i) keep[i] = keep32[i] != 0;
    const Compaction cpt = compact_blocks(nr, keep);
    ROperators out = nr;
    out.beginblock = cpt.beginblock;
    out.op_offset = cpt.op_offset;
    const int64_t ncomp = cpt.op_offset[nr.nrops];
    out.tel.assign(ncomp, 0.0);
    if (ncomp) {
        DevBuf<int64_t> cs, cd, cn;
        auto dcomp = std::make_shared<DevBuf<double>>(ncomp);
        cs.upload(cpt.src, s);
        cd.upload(cpt.dst, s);
        cn.upload(cpt.n, s);
        hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid((int64_t)cpt.n.size())), dim3(kThreads), 0,
                           s, cs.get(), cd.get(), cn.get(), (int)cpt.n.size(), dnew.get(),
                           dcomp->get());
        T3_HIP_CHECK(hipMemcpyAsync(out.tel.data(), dcomp->get(), ncomp * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
        if (keep_dev) *keep_dev = std::move(dcomp);
    }
    T3_HIP_CHECK(hipGetLastError());
    return out;
}

}  // namespace t3

// ---------------------------------------------------------------------------
// update_rOperators_branching: U[dst] = sum_t pref_t * A_t . (O1_t x O2_t) . T_t,
// then the instruction sums of the unique operators.
// ---------------------------------------------------------------------------
namespace t3 {

namespace {

// Y = X x_leg O: X has dims n[3] (column-major), O is (e x n[leg]) column-
// major; Y has n[leg] replaced by e.
void mode_tiles(std::vector<GemmTile>& out, int64_t x, const int n[3], int leg, int64_t o, int e,
                int64_t y, double* flops) {
    GemmTile g;
    if (leg == 0) {
        g.p = Strided{o, 1, e};
        g.q = Strided{x, 1, n[0]};
        g.M = e;
        g.N = n[1] * n[2];
        g.K = n[0];
        g.x_off = y;
        g.ldx = e;
        for (int a = 0; a < tiles_of(g.M); ++a)
            for (int b = 0; b < tiles_of(g.N); ++b) {
                g.ti = a;
                g.tj = b;
                out.push_back(g);
            }
    } else if (leg == 2) {
        g.p = Strided{x, 1, n[0] * n[1]};
        g.q = Strided{o, e, 1};
        g.M = n[0] * n[1];
        g.N = e;
        g.K = n[2];
        g.x_off = y;
        g.ldx = g.M;
        for (int a = 0; a < tiles_of(g.M); ++a)
            for (int b = 0; b < tiles_of(g.N); ++b) {
                g.ti = a;
                g.tj = b;
                out.push_back(g);
            }
    } else {
        // the n[2] slices as one product: row (i, l) at i + l n0 n1 of the
        // input and i + l n0 e of the output
        g.p = Strided{x, 1, n[0]};
        g.p.lo = n[0];
        g.p.hi = (int64_t)n[0] * n[1];
        g.q = Strided{o, e, 1};
        g.M = n[0] * n[2];
        g.N = e;
        g.K = n[1];
        g.x_off = y;
        g.ldx = n[0];
        g.xhi = (int64_t)n[0] * e;
        for (int a = 0; a < tiles_of(g.M); ++a)
            for (int b = 0; b < tiles_of(g.N); ++b) {
                g.ti = a;
                g.tj = b;
                out.push_back(g);
            }
    }
    *flops += 2.0 * e * (double)n[0] * n[1] * n[2];
}

struct KeyHash {
    size_t operator()(const std::array<int64_t, 3>& k) const {
        return std::hash<int64_t>()(k[0] * 1000003 + k[1] * 10007 + k[2]);
    }
};

// dst (n0, n2, n1) = src (n0, n1, n2) with legs 1 and 2 swapped
struct Swap12 {
    int64_t src, dst;
    int n0, n1, n2, pad;
};

__global__ void __launch_bounds__(256)
    swap12_kernel(const Swap12* __restrict__ items, int n, double* arena) {
    for (int it = blockIdx.x; it < n; it += gridDim.x) {
        const Swap12 w = items[it];
        const int64_t tot = (int64_t)w.n0 * w.n1 * w.n2;
        for (int64_t e = threadIdx.x; e < tot; e += blockDim.x) {
            const int64_t i0 = e % w.n0, i2 = (e / w.n0) % w.n2, i1 = e / ((int64_t)w.n0 * w.n2);
            arena[w.dst + e] = arena[w.src + i0 + i1 * w.n0 + i2 * (int64_t)w.n0 * w.n1];
        }
    }
}

}  // namespace

ROperators run_bupdate(const BUpdatePlan& plan, const SiteTensor& T, const ROperators& ops0,
                       const ROperators& ops1, hipStream_t s, RenormWork* work, const double* dev0,
                       const double* dev1, DevArena* keep) {
    const int nl = plan.n_leg, xl = plan.x_leg, yl = plan.y_leg;
    const int64_t nT = T.size(), n0 = ops0.total_size(), n1 = ops1.total_size();
    const int64_t nU = plan.uniq.total_size();
    // Arena: [T | ops0 | ops1 | unique operators | workspace]
    const int64_t o0_base = nT, o1_base = nT + n0, u_base = nT + n0 + n1, w_base = u_base + nU;
    // workspace per chunk of destinations
    constexpr int64_t kWork = (int64_t)1 << 28;
    struct Chunk {
        std::vector<GemmTile> p1, p2;
        std::vector<Swap12> swaps;  // new operators on leg 1: K made contiguous
        std::vector<SumTile> sums;
        std::vector<SumTask> tasks;
        int ndot = 0;      // small-output tiles, first in sums
        int64_t wtop = 0;  // workspace used
    };
    // Tile lists are built on host threads over ranges of destinations of
    // about equal task counts, each range cut into its own chunks (a chunk
    // shares first- and second-stage products among its tasks); the chunks
    // run in range order.  The sums per destination keep their task order,
    // so the results do not depend on the ranges.
    const size_t ndest = plan.dests.size();
    const int nranges = (int)std::min<size_t>(std::max<size_t>(ndest, 1), 64);
    std::vector<size_t> rbeg(nranges + 1, ndest);
    {
        const double per = (double)plan.tasks.size() / nranges;
        size_t d = 0;
        for (int r = 0; r < nranges; ++r) {
            rbeg[r] = d;
            const double target = per * (r + 1);
            while (d < ndest && (double)plan.dests[d].task_end <= target) ++d;
        }
        rbeg[0] = 0;
    }
    std::vector<std::vector<Chunk>> rchunks(nranges);
    std::vector<double> rflops(nranges, 0.0);
    std::vector<int64_t> rwmax(nranges, 0);
    auto build_range = [&](int r) {
        std::vector<Chunk>& chunks = rchunks[r];
        double& flops = rflops[r];
        int64_t& wmax = rwmax[r];
        size_t di = rbeg[r];
        const size_t dend = rbeg[r + 1];
        while (di < dend) {
            Chunk c;
            std::unordered_map<std::array<int64_t, 3>, int64_t, KeyHash> w1, w2, sw;
            int64_t wtop = 0;
            for (; di < dend; ++di) {
                const BUpdateDest& dd = plan.dests[di];
                // workspace this destination needs at most
                int64_t need = 0;
                for (int t = dd.task_begin; t < dd.task_end; ++t) {
                    const BUpdateTask& tk = plan.tasks[t];
                    int64_t s1 = 1, s2 = 1;
                    for (int i = 0; i < 3; ++i) {
                        s1 *= i == xl ? tk.e[i] : tk.d[i];
                        s2 *= i == nl ? tk.d[i] : tk.e[i];
                    }
                    need += s1 + s2;
                    if (nl == 1) need += s2 + (int64_t)tk.e[0] * tk.e[1] * tk.e[2];
                }
                if (need > kWork) throw std::runtime_error("bUpdate: block needs more than 2 GiB of workspace");
                if (wtop + need > kWork && !c.sums.empty()) break;
                SumTile st;
                st.u_off = u_base + dd.dst;
                st.M = dd.M;
                st.N = dd.N;
                st.task_begin = (int)c.tasks.size();
                st.flat = 0;
                st.pad = 0;
                for (int t = dd.task_begin; t < dd.task_end; ++t) {
                    const BUpdateTask& tk = plan.tasks[t];
                    // stage 1: W1 = T x_x O1 (shared by every task with the same T and O1 block)
                    int n1d[3] = {tk.d[0], tk.d[1], tk.d[2]};
                    const std::array<int64_t, 3> k1 = {tk.t, tk.o1, -1};
                    auto it1 = w1.find(k1);
                    int64_t w1off;
                    if (it1 == w1.end()) {
                        w1off = w_base + wtop;
                        int64_t sz = (int64_t)tk.e[xl];
                        for (int i = 0; i < 3; ++i)
                            if (i != xl) sz *= tk.d[i];
                        wtop += sz;
                        mode_tiles(c.p1, tk.t, n1d, xl, o0_base + tk.o1, tk.e[xl], w1off, &flops);
                        w1.emplace(k1, w1off);
                    } else {
                        w1off = it1->second;
                    }
                    // stage 2: W2 = W1 x_y O2
                    int n2d[3] = {tk.d[0], tk.d[1], tk.d[2]};
                    n2d[xl] = tk.e[xl];
                    const std::array<int64_t, 3> k2 = {tk.t, tk.o1, tk.o2};
                    auto it2 = w2.find(k2);
                    int64_t w2off;
                    if (it2 == w2.end()) {
                        w2off = w_base + wtop;
                        int64_t sz = (int64_t)tk.e[yl];
                        for (int i = 0; i < 3; ++i)
                            if (i != yl) sz *= n2d[i];
                        wtop += sz;
                        mode_tiles(c.p2, w1off, n2d, yl, o1_base + tk.o2, tk.e[yl], w2off, &flops);
                        w2.emplace(k2, w2off);
                    } else {
                        w2off = it2->second;
                    }
                    // stage 3: U += pref * A . W2 over the two legs of the old operators
                    const int* e = tk.e;
                    const int* d = tk.d;
                    SumTask sk;
                    sk.pad = 0;
                    sk.pref = tk.pref;
                    if (nl == 2) {
                        sk.r = Strided{tk.a, e[0] * e[1], 1};
                        sk.s = Strided{w2off, 1, e[0] * e[1]};
                        sk.K = e[0] * e[1];
                        c.tasks.push_back(sk);
                    } else if (nl == 0) {
                        sk.r = Strided{tk.a, 1, e[0]};
                        sk.s = Strided{w2off, d[0], 1};
                        sk.K = e[1] * e[2];
                        c.tasks.push_back(sk);
                    } else {
                        // U = sum_g A_g^T W2_g over the last leg: with legs 1 and 2
                        // of A and W2 swapped, one product with K = e0 * e2
                        auto swapped = [&](int64_t src, int n1, int64_t key2) {
                            const std::array<int64_t, 3> key = {src, key2, -2};
                            auto it = sw.find(key);
                            if (it != sw.end()) return it->second;
                            const int64_t off = w_base + wtop;
                            wtop += (int64_t)e[0] * n1 * e[2];
                            c.swaps.push_back(Swap12{src, off, e[0], n1, e[2], 0});
                            sw.emplace(key, off);
                            return off;
                        };
                        const int64_t a2 = swapped(tk.a, e[1], -1);
                        const int64_t w2s = swapped(w2off, d[1], -3);
                        sk.r = Strided{a2, e[0] * e[2], 1};
                        sk.s = Strided{w2s, 1, e[0] * e[2]};
                        sk.K = e[0] * e[2];
                        c.tasks.push_back(sk);
                    }
                    flops += 2.0 * (double)e[0] * e[1] * e[2] * d[nl];
                }
                st.task_end = (int)c.tasks.size();
                for (int a = 0; a < tiles_of(st.M); ++a)
                    for (int b = 0; b < tiles_of(st.N); ++b) {
                        st.ti = a;
                        st.tj = b;
                        c.sums.push_back(st);
                    }
            }
            wmax = std::max(wmax, wtop);
            c.wtop = wtop;
            chunks.push_back(std::move(c));
        }
    };
    run_blocks_threaded(nranges, build_range, (double)plan.tasks.size(), 1e5);
    double flops = 0;
    int64_t wmax = 0;
    for (int r = 0; r < nranges; ++r) {
        flops += rflops[r];
        wmax = std::max(wmax, rwmax[r]);
    }
    // Consecutive chunks run as one while their workspaces fit together in
    // max(largest chunk's, 8 Mi values): the ranges exist for the host
    // threads, and every device chunk costs its list uploads, launches and a
    // synchronisation.  A chunk appended to another has its workspace
    // offsets (those at or past w_base) moved past the other's workspace and
    // its task indices past the other's tasks; every destination keeps its
    // tasks in order, so the sums are unchanged.
    const int64_t wcap = std::max<int64_t>(wmax, (int64_t)1 << 23);
    std::vector<Chunk> chunks;
    for (int r = 0; r < nranges; ++r)
        for (Chunk& c : rchunks[r]) {
            if (chunks.empty() || chunks.back().wtop + c.wtop > wcap) {
                chunks.push_back(std::move(c));
                continue;
            }
            Chunk& m = chunks.back();
            const int64_t dw = m.wtop;
            const int dt = (int)m.tasks.size();
            auto rel = [&](int64_t& off) {
                if (off >= w_base) off += dw;
            };
            for (GemmTile g : c.p1) {
                rel(g.p.off);
                rel(g.q.off);
                rel(g.x_off);
                m.p1.push_back(g);
            }
            for (GemmTile g : c.p2) {
                rel(g.p.off);
                rel(g.q.off);
                rel(g.x_off);
                m.p2.push_back(g);
            }
            for (Swap12 w : c.swaps) {
                rel(w.src);
                rel(w.dst);
                m.swaps.push_back(w);
            }
            for (SumTask t : c.tasks) {
                rel(t.r.off);
                rel(t.s.off);
                m.tasks.push_back(t);
            }
            for (SumTile st : c.sums) {
                st.task_begin += dt;
                st.task_end += dt;
                m.sums.push_back(st);
            }
            m.wtop += c.wtop;
        }
    for (Chunk& c : chunks) {
        wmax = std::max(wmax, c.wtop);
        c.ndot = partition_sums(c.sums);
    }

    DevBuf<double> arena(std::max<int64_t>(1, w_base + wmax));
    if (nT) T3_HIP_CHECK(hipMemcpyAsync(arena.get(), T.tel.data(), nT * sizeof(double),
                                        hipMemcpyHostToDevice, s));
    if (n0)
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + o0_base, dev0 ? dev0 : ops0.tel.data(),
                                    n0 * sizeof(double),
                                    dev0 ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (n1)
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + o1_base, dev1 ? dev1 : ops1.tel.data(),
                                    n1 * sizeof(double),
                                    dev1 ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_zero<double>, dim3(grid((nU + 255) / 256)), dim3(256), 0, s,
                       arena.get() + u_base, nU);
    DevBuf<unsigned int> dkeep(1);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    // device tile lists, grown to the largest chunk and reused (the
    // per-chunk synchronisation below keeps an upload from overwriting lists
    // still in use)
    DevBuf<GemmTile> d1, d2;
    DevBuf<SumTile> ds;
    DevBuf<SumTask> dt;
    DevBuf<Swap12> dw;
    for (const Chunk& c : chunks) {
        d1.upload(c.p1, s);
        d2.upload(c.p2, s);
        ds.upload(c.sums, s);  // small-output tiles first (partitioned at chunk close)
        dt.upload(c.tasks, s);
        if (!c.p1.empty())
            hipLaunchKernelGGL(tile::gemm_tiles_kernel, dim3(grid((int64_t)c.p1.size())), dim3(256), 0,
                               s, d1.get(), (int)c.p1.size(), arena.get());
        if (!c.p2.empty())
            hipLaunchKernelGGL(tile::gemm_tiles_kernel, dim3(grid((int64_t)c.p2.size())), dim3(256), 0,
                               s, d2.get(), (int)c.p2.size(), arena.get());
        dw.upload(c.swaps, s);
        if (!c.swaps.empty())
            hipLaunchKernelGGL(swap12_kernel, dim3(grid((int64_t)c.swaps.size())), dim3(256), 0, s,
                               dw.get(), (int)c.swaps.size(), arena.get());
        launch_sums(ds.get(), c.ndot, (int)c.sums.size(), dt.get(), arena.get(), arena.get(),
                    dkeep.get(), s);
        // the next chunk's uploads reuse the lists: wait for the kernels
        T3_HIP_CHECK(hipStreamSynchronize(s));
    }
    T3_HIP_CHECK(hipEventRecord(e1, s));
    T3_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    T3_HIP_CHECK(hipGetLastError());
    if (work) {
        work->flops = flops;
        work->bytes = 8.0 * (double)(nT + n0 + n1 + plan.nr.total_size());
        work->kernel_ms = ms;
    }
    return finish_sum(plan.nr, plan.groups, plan.stasks, arena.get() + u_base, s, nullptr, keep);
}

}  // namespace t3
