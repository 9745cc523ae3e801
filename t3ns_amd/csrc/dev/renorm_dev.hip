// Device kernels of the renormalisation plans: grouped scaled sums of
// operator blocks, empty-block detection and compaction.  All are
// bandwidth-bound (each element is read once per contributing term and
// written once).
#include "renorm_dev.hpp"

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
    const int64_t nold = old.total_size(), nnew = nr.total_size();
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
