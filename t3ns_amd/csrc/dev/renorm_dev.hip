// Device kernels of the renormalisation plans: grouped scaled sums of
// operator blocks, empty-block detection and compaction.  All are
// bandwidth-bound (each element is read once per contributing term and
// written once).
#include "renorm_dev.hpp"
#include "tile_gemm.hpp"

#include <algorithm>

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

ROperators run_append(const AppendPlan& plan, const ROperators& old, hipStream_t s) {
    const ROperators& nr = plan.nr;
    const int64_t nnew = nr.total_size();
    // flat block index of every (op, block)
    std::vector<int> flat_of_op(nr.nrops + 1, 0);
    for (int k = 0; k < nr.nrops; ++k) flat_of_op[k + 1] = flat_of_op[k] + nr.nblocks_op(k);
    const int nflat = flat_of_op[nr.nrops];
    std::vector<DevGroup> hg(plan.groups.size());
    for (size_t i = 0; i < plan.groups.size(); ++i) {
        const AppendGroup& g = plan.groups[i];
        hg[i] = DevGroup{g.dst, g.n, g.task_begin, g.task_end, flat_of_op[g.op] + g.block, 0};
    }
    std::vector<int64_t> tsrc(plan.tasks.size());
    std::vector<double> tfac(plan.tasks.size()), tpref(plan.tasks.size());
    for (size_t i = 0; i < plan.tasks.size(); ++i) {
        tsrc[i] = plan.tasks[i].src;
        tfac[i] = plan.tasks[i].factor;
        tpref[i] = plan.tasks[i].pref;
    }
    DevBuf<DevGroup> dg;
    DevBuf<int64_t> dsrc;
    DevBuf<double> dfac, dpref, dold, dnew;
    DevBuf<unsigned char> dkeep;
    dg.upload(hg, s);
    dsrc.upload(tsrc, s);
    dfac.upload(tfac, s);
    dpref.upload(tpref, s);
    dold.upload(old.tel, s);
    dnew.alloc(std::max<int64_t>(1, nnew));
    dkeep.alloc(std::max(1, nflat));
    hipLaunchKernelGGL(fill_zero<unsigned char>, dim3(grid((nflat + 255) / 256)), dim3(256), 0, s,
                       dkeep.get(), (int64_t)nflat);
    if (!hg.empty())
        hipLaunchKernelGGL(append_kernel, dim3(grid((int64_t)hg.size())), dim3(kThreads), 0, s,
                           dg.get(), (int)hg.size(), dsrc.get(), dfac.get(), dpref.get(), dold.get(),
                           dnew.get(), dkeep.get());
    std::vector<unsigned char> keep(nflat);
    if (nflat)
        T3_HIP_CHECK(hipMemcpyAsync(keep.data(), dkeep.get(), nflat, hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    const Compaction cpt = compact_blocks(nr, keep);
    ROperators out = nr;
    out.beginblock = cpt.beginblock;
    out.op_offset = cpt.op_offset;
    const int64_t ncomp = cpt.op_offset[nr.nrops];
    out.tel.assign(ncomp, 0.0);
    if (ncomp) {
        DevBuf<int64_t> cs, cd, cn;
        DevBuf<double> dcomp(ncomp);
        cs.upload(cpt.src, s);
        cd.upload(cpt.dst, s);
        cn.upload(cpt.n, s);
        hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid((int64_t)cpt.n.size())), dim3(kThreads), 0,
                           s, cs.get(), cd.get(), cn.get(), (int)cpt.n.size(), dnew.get(), dcomp.get());
        T3_HIP_CHECK(hipMemcpyAsync(out.tel.data(), dcomp.get(), ncomp * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
    }
    T3_HIP_CHECK(hipGetLastError());
    return out;
}

}  // namespace t3

// ---------------------------------------------------------------------------
// update_rOperators_physical: U[dst] = sum_t pref_t * A_t O_t B_t.
// ---------------------------------------------------------------------------
namespace t3 {

namespace {

using tile::GemmTile;
using tile::kKStep;
using tile::kTile;
using tile::Strided;
using tile::tile_slab;
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

__global__ void __launch_bounds__(256)
    pupdate_sum_kernel(const SumTile* __restrict__ tiles, int ntiles,
                       const SumTask* __restrict__ tasks, const double* __restrict__ base,
                       double* __restrict__ out, unsigned int* __restrict__ keep) {
    __shared__ double Ps[kKStep][kTile + 1];
    __shared__ double Qs[kKStep][kTile + 1];
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const SumTile g = tiles[it];
        const int i0 = g.ti * kTile, j0 = g.tj * kTile;
        double u[4][4] = {};
        for (int t = g.task_begin; t < g.task_end; ++t) {
            const SumTask tk = tasks[t];
            double acc[4][4] = {};
            for (int k0 = 0; k0 < tk.K; k0 += kKStep)
                tile_slab(base, tk.r, tk.s, g.M, g.N, tk.K, i0, j0, k0, Ps, Qs, acc);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) u[a][b] = fma(tk.pref, acc[a][b], u[a][b]);
        }
        const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
        int any = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + ty + 16 * b;
            if (j >= g.N) continue;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = i0 + tx + 16 * a;
                if (i < g.M) {
                    out[g.u_off + i + (int64_t)j * g.M] = u[a][b];
                    any |= fabs(u[a][b]) >= 1e-20;
                }
            }
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0 && any) atomicOr(&keep[g.flat], 1u);
    }
}

}  // namespace

ROperators run_pupdate(const PUpdatePlan& plan, const ROperators& pops, const SiteTensor& T,
                       hipStream_t s, RenormWork* work) {
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
        T3_HIP_CHECK(hipMemcpyAsync(arena.get() + oldoff, pops.tel.data(), nold * sizeof(double),
                                    hipMemcpyHostToDevice, s));
    DevBuf<GemmTile> dfirst;
    DevBuf<SumTile> dsums;
    DevBuf<SumTask> dtasks;
    DevBuf<double> dnew(std::max<int64_t>(1, nnew));
    DevBuf<unsigned int> dkeep(std::max(1, nflat));
    dfirst.upload(first, s);
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
    if (!sums.empty())
        hipLaunchKernelGGL(pupdate_sum_kernel, dim3(grid((int64_t)sums.size())), dim3(256), 0, s,
                           dsums.get(), (int)sums.size(), dtasks.get(), arena.get(), dnew.get(),
                           dkeep.get());
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
        DevBuf<double> dcomp(ncomp);
        cs.upload(cpt.src, s);
        cd.upload(cpt.dst, s);
        cn.upload(cpt.n, s);
        hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid((int64_t)cpt.n.size())), dim3(kThreads), 0,
                           s, cs.get(), cd.get(), cn.get(), (int)cpt.n.size(), dnew.get(), dcomp.get());
        T3_HIP_CHECK(hipMemcpyAsync(out.tel.data(), dcomp.get(), ncomp * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        T3_HIP_CHECK(hipStreamSynchronize(s));
    }
    T3_HIP_CHECK(hipGetLastError());
    return out;
}

}  // namespace t3
