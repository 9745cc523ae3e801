// Device effective Hamiltonian: work-list construction and kernels.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "heff_dev.hpp"

namespace t3 {

namespace {
constexpr int kThreads = 256;
__constant__ int kOrd[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
const int hOrd[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// out = O x_p in  (or out += alpha * O x_p in): mode-p product of a
// three-leg block (memory dims a0,a1,a2, first fastest) with a
// column-major (dn x a_p) operator block.
template <typename T>
__device__ __forceinline__ void mode_product(const T* __restrict__ in, T* __restrict__ out,
                                             const T* __restrict__ O, int a0, int a1, int a2,
                                             int p, int dn, T alpha, bool acc) {
    int o0 = a0, o1 = a1, o2 = a2;
    if (p == 0) o0 = dn;
    else if (p == 1) o1 = dn;
    else o2 = dn;
    const int total = o0 * o1 * o2;
    const int s1 = a0, s2 = a0 * a1;
    const int K = p == 0 ? a0 : (p == 1 ? a1 : a2);
    const int st = p == 0 ? 1 : (p == 1 ? s1 : s2);
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int i0 = idx % o0;
        const int r = idx / o0;
        const int i1 = r % o1;
        const int i2 = r / o1;
        int ip;
        const T* src;
        if (p == 0) {
            ip = i0;
            src = in + i1 * s1 + i2 * s2;
        } else if (p == 1) {
            ip = i1;
            src = in + i0 + i2 * s2;
        } else {
            ip = i2;
            src = in + i0 + i1 * s1;
        }
        T s = 0;
        for (int k = 0; k < K; ++k) s += O[ip + k * dn] * src[k * st];
        if (acc) out[idx] += alpha * s;
        else out[idx] = s;
    }
}

// One workgroup owns one output block at a time and accumulates every
// contributing term into it: no atomics, deterministic summation order.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    heff_matvec_kernel(const DevBlock* __restrict__ blocks, int nblocks,
                       const DevPair* __restrict__ pairs, const DevTerm* __restrict__ terms,
                       const T* __restrict__ op0, const T* __restrict__ op1,
                       const T* __restrict__ op2, const T* __restrict__ x, T* __restrict__ y,
                       T* __restrict__ scratch, int64_t scratch_stride, int isdmrg, int m0, int m1,
                       int m2) {
    T* w1base = scratch + (int64_t)blockIdx.x * scratch_stride;
    const T* opb[3] = {op0, op1, op2};
    const int map[3] = {m0, m1, m2};
    int pos[3];
    pos[m0] = 0;
    pos[m1] = 1;
    pos[m2] = 2;
    for (int bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const DevBlock B = blocks[bi];
        T* Y = y + B.y_off;
        for (int pi = B.pair_begin; pi < B.pair_end; ++pi) {
            const DevPair P = pairs[pi];
            const T* X = x + P.x_off;
            const int nst = isdmrg ? 2 : 3;
            // memory-order dims before each stage
            int dm[3] = {P.d[map[0]], P.d[map[1]], P.d[map[2]]};
            int dims_after1[3], sz1 = 0;
            {
                const int o = isdmrg ? (P.order == 0 ? 0 : 1) : kOrd[P.order][0];
                for (int k = 0; k < 3; ++k) dims_after1[k] = dm[k];
                dims_after1[pos[o]] = B.d[o];
                sz1 = dims_after1[0] * dims_after1[1] * dims_after1[2];
            }
            T* w1 = w1base;
            T* w2 = w1base + sz1;
            for (int ti = P.term_begin; ti < P.term_end; ++ti) {
                const DevTerm tm = terms[ti];
                const T alpha = (T)tm.alpha;
                int cur[3] = {dm[0], dm[1], dm[2]};
                const T* src = X;
                for (int s = 0; s < nst; ++s) {
                    const int o = isdmrg ? (P.order == 0 ? s : 1 - s) : kOrd[P.order][s];
                    const int p = pos[o];
                    const T* O = opb[o] + tm.op[o];
                    const bool last = s == nst - 1;
                    T* dst = last ? Y : (s == 0 ? w1 : w2);
                    mode_product<T>(src, dst, O, cur[0], cur[1], cur[2], p, B.d[o], alpha, last);
                    __syncthreads();
                    cur[p] = B.d[o];
                    src = dst;
                }
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    heff_diag_kernel(const DevDiagBlock* __restrict__ blocks, int nblocks,
                     const DevTerm* __restrict__ terms, const T* __restrict__ op0,
                     const T* __restrict__ op1, const T* __restrict__ op2, double* __restrict__ diag,
                     int isdmrg, int m0, int m1, int m2) {
    const T* opb[3] = {op0, op1, op2};
    for (int bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const DevDiagBlock B = blocks[bi];
        const int M = B.d[m0], N = B.d[m1], K = isdmrg ? 1 : B.d[m2];
        const int total = M * N * K;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int m = idx % M;
            const int n = (idx / M) % N;
            const int k = idx / (M * N);
            double acc = 0;
            for (int ti = B.term_begin; ti < B.term_end; ++ti) {
                const DevTerm tm = terms[ti];
                double v = tm.alpha * (double)opb[m0][tm.op[m0] + (int64_t)m * (M + 1)] *
                           (double)opb[m1][tm.op[m1] + (int64_t)n * (N + 1)];
                if (!isdmrg) v *= (double)opb[m2][tm.op[m2] + (int64_t)k * (K + 1)];
                acc += v;
            }
            diag[B.off + idx] += acc;
        }
    }
}

template <typename T>
__global__ void fill_kernel(T* p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename A, typename B>
__global__ void convert_kernel(const A* __restrict__ src, B* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (B)src[i];
}

inline int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}
}  // namespace

HeffDevice::HeffDevice(const HeffInput& in, int precision, hipStream_t stream)
    : prec_(precision) {
    plan_ = build_heff_plan(in);
    isdmrg_ = plan_.isdmrg;
    for (int k = 0; k < 3; ++k) map_[k] = plan_.map[k];
    size_ = in.site->size();
    const int nops = isdmrg_ ? 2 : 3;
    const InstructionSet& is = *in.iset;

    // Flatten the plan into work lists.
    std::vector<DevPair> hp;
    std::vector<DevTerm> ht;
    std::vector<DevBlock> hb;
    std::vector<double> cost(plan_.nrblocks, 0.0);
    int64_t maxscratch = 1;
    for (int sb = 0; sb < plan_.nrblocks; ++sb) {
        DevBlock B{};
        B.y_off = plan_.beginblock[sb];
        for (int k = 0; k < 3; ++k) B.d[k] = plan_.dims[sb][k];
        B.pair_begin = (int)hp.size();
        for (int pi = plan_.pair_begin[sb]; pi < plan_.pair_begin[sb + 1]; ++pi) {
            const PlanPair& pp = plan_.pairs[pi];
            DevPair P{};
            P.x_off = plan_.beginblock[pp.oldsb];
            for (int k = 0; k < 3; ++k) P.d[k] = plan_.dims[pp.oldsb][k];
            P.order = pp.bestorder;
            P.term_begin = (int)ht.size();
            for (const PlanEntry& e : pp.entries) {
                for (int i = is.MPOc_beg[e.mpo_group]; i < is.MPOc_beg[e.mpo_group + 1]; ++i) {
                    const Instruction& ins = is.instr[i];
                    DevTerm t{};
                    bool ok = true;
                    for (int k = 0; k < nops; ++k) {
                        const ROperators& op = *in.ops[k];
                        const int o = ins.instr[k];
                        if (op.block_size(o, e.sbops[k]) == 0) {
                            ok = false;
                            break;
                        }
                        t.op[k] = op.op_offset[o] + op.beginblock[o][e.sbops[k]];
                    }
                    if (!ok) continue;
                    t.alpha = ins.pref * e.prefactor;
                    ht.push_back(t);
                }
            }
            P.term_end = (int)ht.size();
            if (P.term_end == P.term_begin) continue;
            // scratch: stage-1 and stage-2 outputs
            int w[3] = {P.d[0], P.d[1], P.d[2]};
            int64_t need = 0;
            for (int s = 0; s < (isdmrg_ ? 1 : 2); ++s) {
                const int o = isdmrg_ ? (P.order == 0 ? 0 : 1) : hOrd[P.order][s];
                w[o] = B.d[o];
                need += (int64_t)w[0] * w[1] * w[2];
            }
            maxscratch = std::max(maxscratch, need);
            cost[sb] += (P.term_end - P.term_begin) *
                        heff_term_flops(P.d, B.d, map_, P.order, isdmrg_);
            hp.push_back(P);
        }
        B.pair_end = (int)hp.size();
        hb.push_back(B);
    }
    // Largest blocks first so the persistent grid drains evenly.
    std::vector<int> idx(hb.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<DevBlock> hbs;
    for (int i : idx)
        if (hb[i].pair_end > hb[i].pair_begin) hbs.push_back(hb[i]);

    // Diagonal work list.
    std::vector<DiagEntry> de = build_heff_diag_plan(in, plan_);
    std::vector<DevDiagBlock> hdb;
    std::vector<DevTerm> hdt;
    for (size_t i = 0; i < de.size();) {
        const int sb = de[i].sb;
        DevDiagBlock D{};
        D.off = plan_.beginblock[sb];
        for (int k = 0; k < 3; ++k) D.d[k] = plan_.dims[sb][k];
        D.term_begin = (int)hdt.size();
        for (; i < de.size() && de[i].sb == sb; ++i) {
            const DiagEntry& e = de[i];
            for (int j = is.MPOc_beg[e.mpo_group]; j < is.MPOc_beg[e.mpo_group + 1]; ++j) {
                const Instruction& ins = is.instr[j];
                DevTerm t{};
                bool ok = true;
                for (int k = 0; k < nops; ++k) {
                    const ROperators& op = *in.ops[k];
                    const int o = ins.instr[k];
                    if (op.block_size(o, e.sbops[k]) == 0) {
                        ok = false;
                        break;
                    }
                    t.op[k] = op.op_offset[o] + op.beginblock[o][e.sbops[k]];
                }
                if (!ok) continue;
                t.alpha = ins.pref * e.prefactor;
                hdt.push_back(t);
            }
        }
        D.term_end = (int)hdt.size();
        if (D.term_end > D.term_begin) hdb.push_back(D);
    }
    ndblocks_ = (int)hdb.size();

    stats_.nblocks = (int64_t)hbs.size();
    stats_.npairs = (int64_t)hp.size();
    stats_.nterms = (int64_t)ht.size();
    stats_.size = size_;
    stats_.flops = plan_.flops;
    stats_.scratch_per_wg = maxscratch;

    blocks_.upload(hbs, stream);
    pairs_.upload(hp, stream);
    terms_.upload(ht, stream);
    dblocks_.upload(hdb, stream);
    dterms_.upload(hdt, stream);
    for (int k = 0; k < nops; ++k) {
        const ROperators& op = *in.ops[k];
        stats_.op_elems[k] = op.total_size();
        if ((int64_t)op.tel.size() != op.total_size())
            throw std::runtime_error("HeffDevice: operator host data missing");
        ops64_[k].upload(op.tel, stream);
    }
    if (prec_ == PREC_FP32) {
        for (int k = 0; k < nops; ++k) {
            const int64_t n = stats_.op_elems[k];
            ops32_[k].alloc(n);
            if (n)
                hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0,
                                   stream, ops64_[k].get(), ops32_[k].get(), n);
        }
        x32_.alloc(size_);
        y32_.alloc(size_);
    }
    nwg_ = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)hbs.size(), 4096));
    if (prec_ == PREC_FP32) scratch32_.alloc((size_t)nwg_ * maxscratch);
    else scratch64_.alloc((size_t)nwg_ * maxscratch);
    T3_HIP_CHECK(hipGetLastError());
    T3_HIP_CHECK(hipStreamSynchronize(stream));
}

HeffDevice::~HeffDevice() = default;

void HeffDevice::matvec(const double* dvec, double* dres, hipStream_t s) {
    const int nb = (int)stats_.nblocks;
    if (prec_ == PREC_FP64) {
        hipLaunchKernelGGL((fill_kernel<double>), dim3(grid_for(size_)), dim3(256), 0, s, dres,
                           size_, 0.0);
        if (nb)
            hipLaunchKernelGGL((heff_matvec_kernel<double>), dim3(nwg_), dim3(kThreads), 0, s,
                               blocks_.get(), nb, pairs_.get(), terms_.get(), ops64_[0].get(),
                               ops64_[1].get(), ops64_[2].get(), dvec, dres, scratch64_.get(),
                               stats_.scratch_per_wg, isdmrg_, map_[0], map_[1], map_[2]);
    } else {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(size_)), dim3(256), 0, s,
                           dvec, x32_.get(), size_);
        hipLaunchKernelGGL((fill_kernel<float>), dim3(grid_for(size_)), dim3(256), 0, s,
                           y32_.get(), size_, 0.0f);
        if (nb)
            hipLaunchKernelGGL((heff_matvec_kernel<float>), dim3(nwg_), dim3(kThreads), 0, s,
                               blocks_.get(), nb, pairs_.get(), terms_.get(), ops32_[0].get(),
                               ops32_[1].get(), ops32_[2].get(), x32_.get(), y32_.get(),
                               scratch32_.get(), stats_.scratch_per_wg, isdmrg_, map_[0], map_[1],
                               map_[2]);
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(size_)), dim3(256), 0, s,
                           y32_.get(), dres, size_);
    }
    T3_HIP_CHECK(hipGetLastError());
}

void HeffDevice::diagonal(double* ddiag, hipStream_t s) {
    hipLaunchKernelGGL((fill_kernel<double>), dim3(grid_for(size_)), dim3(256), 0, s, ddiag, size_,
                       0.0);
    if (ndblocks_)
        hipLaunchKernelGGL((heff_diag_kernel<double>), dim3(std::min(ndblocks_, 8192)), dim3(256), 0,
                           s, dblocks_.get(), ndblocks_, dterms_.get(), ops64_[0].get(),
                           ops64_[1].get(), ops64_[2].get(), ddiag, isdmrg_, map_[0], map_[1],
                           map_[2]);
    T3_HIP_CHECK(hipGetLastError());
}

}  // namespace t3
