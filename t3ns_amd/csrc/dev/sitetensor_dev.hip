// Multisite tensor from one-site tensors on the device.
#include "sitetensor_dev.hpp"

#include <algorithm>

#include "tile_gemm.hpp"

namespace t3 {

namespace {

using tile::GemmTile;
using tile::Strided;
using tile::tiles_of;

// Reorders the three leg groups of a block (dims n0, n1, n2, column-major)
// into the order perm[0], perm[1], perm[2].
struct PermItem {
    int64_t src, dst;
    int n[3];
    int pad;
};

__global__ void __launch_bounds__(256)
    permute_groups_kernel(const PermItem* __restrict__ items, int nitems, int p0, int p1, int p2,
                          const double* __restrict__ in, double* __restrict__ out) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const PermItem g = items[it];
        const int64_t st[3] = {1, g.n[0], (int64_t)g.n[0] * g.n[1]};
        const int perm[3] = {p0, p1, p2};
        const int64_t m0 = g.n[perm[0]], m1 = g.n[perm[1]];
        const int64_t total = (int64_t)g.n[0] * g.n[1] * g.n[2];
        for (int64_t d = threadIdx.x; d < total; d += blockDim.x) {
            const int64_t c0 = d % m0, c1 = (d / m0) % m1, c2 = d / (m0 * m1);
            out[g.dst + d] = in[g.src + c0 * st[perm[0]] + c1 * st[perm[1]] + c2 * st[perm[2]]];
        }
    }
}

template <typename T>
__global__ void fill_zero_kernel(T* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

int grid_of(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(16384, n)); }

}  // namespace

std::vector<double> run_makesite(const MakesitePlan& plan, const SiteTensor* const* sites,
                                 hipStream_t s, RenormWork* work) {
    const int ns = plan.nrsites, I = plan.innersite;
    const int nsteps = (int)plan.steps.size();
    const size_t nb = plan.blocks.size();
    // Arena: [one-site tensors | multisite | one work region per step]
    const int64_t out_base = plan.site_arena_size;
    int64_t top = out_base + plan.out_size;
    std::vector<std::vector<GemmTile>> tiles(nsteps);
    std::vector<PermItem> perms;
    double flops = 0;
    for (size_t bi = 0; bi < nb; ++bi) {
        const MakesiteBlock& B = plan.blocks[bi];
        int64_t cur[3] = {B.odim[I][0], B.odim[I][1], B.odim[I][2]};
        int64_t r_off = B.site_off[I];
        for (int t = 0; t < nsteps; ++t) {
            const MakesiteStep& st = plan.steps[t];
            const int j = st.inner_leg, k = st.site_leg;
            const int* od = B.odim[st.site];
            const int64_t A = j == 0 ? 1 : (j == 1 ? cur[0] : cur[0] * cur[1]);
            const int64_t K = cur[j];
            const int64_t Bn = j == 2 ? 1 : (j == 1 ? cur[2] : cur[1] * cur[2]);
            const int64_t E = (int64_t)od[1] * od[k == 0 ? 2 : 0];
            const int64_t s_off = B.site_off[st.site];
            const bool last = t == nsteps - 1;
            int64_t x_off;
            if (last && !plan.permuted) {
                x_off = out_base + B.out_off;
            } else {
                x_off = top;
                top += A * E * Bn;
            }
            // S as (k, e): leg 0 -> S is (K x E); leg 2 -> S is (E x K)
            const Strided S = k == 0 ? Strided{s_off, 1, (int)K} : Strided{s_off, (int)E, 1};
            GemmTile g;
            if (A == 1) {
                // X (E x Bn) = S^T R
                g.p = k == 0 ? Strided{s_off, (int)K, 1} : Strided{s_off, 1, (int)E};
                g.q = Strided{r_off, 1, (int)K};
                g.x_off = x_off;
                g.M = (int)E;
                g.N = (int)Bn;
                g.K = (int)K;
                g.ldx = (int)E;
                for (int a = 0; a < tiles_of(g.M); ++a)
                    for (int b = 0; b < tiles_of(g.N); ++b) {
                        g.ti = a;
                        g.tj = b;
                        tiles[t].push_back(g);
                    }
            } else {
                // per b: X_b (A x E) = R_b S
                for (int64_t b = 0; b < Bn; ++b) {
                    g.p = Strided{r_off + b * A * K, 1, (int)A};
                    g.q = S;
                    g.x_off = x_off + b * A * E;
                    g.M = (int)A;
                    g.N = (int)E;
                    g.K = (int)K;
                    g.ldx = (int)A;
                    for (int a = 0; a < tiles_of(g.M); ++a)
                        for (int c = 0; c < tiles_of(g.N); ++c) {
                            g.ti = a;
                            g.tj = c;
                            tiles[t].push_back(g);
                        }
                }
            }
            flops += 2.0 * (double)A * K * E * Bn;
            cur[j] = E;
            r_off = x_off;
        }
        if (plan.permuted)
            perms.push_back(PermItem{r_off, out_base + B.out_off, {(int)cur[0], (int)cur[1], (int)cur[2]}, 0});
    }
    if (top > ((int64_t)1 << 34)) throw std::runtime_error("makesite: device arena above 128 GiB");

    DevBuf<double> arena(std::max<int64_t>(1, top));
    for (int i = 0; i < ns; ++i)
        if (sites[i]->size())
            T3_HIP_CHECK(hipMemcpyAsync(arena.get() + plan.site_arena_off[i], sites[i]->tel.data(),
                                        sites[i]->size() * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(grid_of((plan.out_size + 255) / 256)), dim3(256),
                       0, s, arena.get() + out_base, plan.out_size);
    std::vector<DevBuf<GemmTile>> dtiles(nsteps);
    for (int t = 0; t < nsteps; ++t) dtiles[t].upload(tiles[t], s);
    DevBuf<PermItem> dperm;
    dperm.upload(perms, s);
    hipEvent_t e0, e1;
    T3_HIP_CHECK(hipEventCreate(&e0));
    T3_HIP_CHECK(hipEventCreate(&e1));
    T3_HIP_CHECK(hipEventRecord(e0, s));
    for (int t = 0; t < nsteps; ++t)
        if (!tiles[t].empty())
            hipLaunchKernelGGL(tile::gemm_tiles_kernel, dim3(grid_of((int64_t)tiles[t].size())),
                               dim3(256), 0, s, dtiles[t].get(), (int)tiles[t].size(), arena.get());
    if (!perms.empty())
        hipLaunchKernelGGL(permute_groups_kernel, dim3(grid_of((int64_t)perms.size())), dim3(256), 0, s,
                           dperm.get(), (int)perms.size(), plan.perm[0], plan.perm[1], plan.perm[2],
                           arena.get(), arena.get());
    T3_HIP_CHECK(hipEventRecord(e1, s));
    std::vector<double> out(plan.out_size);
    if (plan.out_size)
        T3_HIP_CHECK(hipMemcpyAsync(out.data(), arena.get() + out_base, plan.out_size * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    T3_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    T3_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    T3_HIP_CHECK(hipGetLastError());
    if (work) {
        work->flops = flops;
        work->bytes = 8.0 * (double)(plan.site_arena_size + plan.out_size);
        work->kernel_ms = ms;
    }
    return out;
}

}  // namespace t3
