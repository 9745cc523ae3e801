// Tiled FP64 GEMMs over strided operands, shared by the renormalisation and
// site-tensor kernels: every work item is one 64 x 64 output tile of
// X = P Q with P(i, k) and Q(k, j) read through arbitrary strides (so
// transposed and batched operands need no copies), K reduced in slabs of 16
// staged through LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace t3 {
namespace tile {

constexpr int kTile = 64;   // output tile edge
constexpr int kKStep = 16;  // reduction slab

// Strided operand: element (i, k) at base[off + i * si + k * sk].
struct Strided {
    int64_t off;
    int si, sk;
};

// One output tile of X = P Q (M x N, column-major with leading dimension
// ldx at x_off); q is read as (k, j).
struct GemmTile {
    Strided p, q;
    int64_t x_off;
    int M, N, K;
    int ti, tj;
    int ldx;
};

inline int tiles_of(int n) { return (n + kTile - 1) / kTile; }

// One K-slab of a 64 x 64 tile product: loads P(64 x 16) and Q(16 x 64)
// into LDS (zero padded) and accumulates acc[a][b] for rows tx + 16a,
// columns ty + 16b.  The load pattern follows the unit-stride index so that
// neighbouring lanes read neighbouring addresses.
__device__ __forceinline__ void tile_slab(const double* __restrict__ base, const Strided& P,
                                          const Strided& Q, int M, int N, int K, int i0, int j0,
                                          int k0, double (*Ps)[kTile + 1],
                                          double (*Qs)[kTile + 1], double acc[4][4]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int i, k;
        if (P.si == 1) {
            i = t % kTile;
            k = t / kTile + 4 * r;
        } else {
            k = t % kKStep;
            i = t / kKStep + 16 * r;
        }
        const int gi = i0 + i, gk = k0 + k;
        Ps[k][i] = (gi < M && gk < K) ? base[P.off + (int64_t)gi * P.si + (int64_t)gk * P.sk] : 0.0;
        int j, kk;
        if (Q.si == 1) {
            kk = t % kKStep;
            j = t / kKStep + 16 * r;
        } else {
            j = t % kTile;
            kk = t / kTile + 4 * r;
        }
        const int gj = j0 + j, gkk = k0 + kk;
        Qs[kk][j] =
            (gj < N && gkk < K) ? base[Q.off + (int64_t)gkk * Q.si + (int64_t)gj * Q.sk] : 0.0;
    }
    __syncthreads();
    const int tx = t % 16, ty = t / 16;
#pragma unroll 4
    for (int k = 0; k < kKStep; ++k) {
        double pa[4], qb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) pa[a] = Ps[k][tx + 16 * a];
#pragma unroll
        for (int b = 0; b < 4; ++b) qb[b] = Qs[k][ty + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(pa[a], qb[b], acc[a][b]);
    }
    __syncthreads();
}

// X tiles = P Q, all operands in one arena (inputs and outputs of one
// launch must not overlap).  256 threads, grid-stride over the tiles.
// Internal linkage: every translation unit that launches it has its copy.
static __global__ void __launch_bounds__(256)
    gemm_tiles_kernel(const GemmTile* __restrict__ tiles, int ntiles, double* base) {
    __shared__ double Ps[kKStep][kTile + 1];
    __shared__ double Qs[kKStep][kTile + 1];
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const GemmTile g = tiles[it];
        const int i0 = g.ti * kTile, j0 = g.tj * kTile;
        double acc[4][4] = {};
        for (int k0 = 0; k0 < g.K; k0 += kKStep)
            tile_slab(base, g.p, g.q, g.M, g.N, g.K, i0, j0, k0, Ps, Qs, acc);
        const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + ty + 16 * b;
            if (j >= g.N) continue;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = i0 + tx + 16 * a;
                if (i < g.M) base[g.x_off + i + (int64_t)j * g.ldx] = acc[a][b];
            }
        }
    }
}

}  // namespace tile
}  // namespace t3
