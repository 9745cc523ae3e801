// Tiled FP64 GEMMs over strided operands, shared by the renormalisation and
// site-tensor kernels: every work item is one 64 x 64 output tile of
// X = P Q with P(i, k) and Q(k, j) read through arbitrary strides (so
// transposed and batched operands need no copies), computed on the matrix
// cores as four 16-row strips (strips and column blocks past the matrix
// edge are skipped, so small blocks pay 16-granular padding only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace t3 {
namespace tile {

constexpr int kTile = 64;  // output tile edge

// Strided operand: element (i, k) at base[off + i * si + k * sk].
struct Strided {
    int64_t off;
    int si, sk;
    // optional two-level first index (batched products folded into one):
    // i -> (i % lo) si + (i / lo) hi; lo = 0: i si
    int lo = 0;
    int64_t hi = 0;
};
// FIB: the two-level form is in use (a template argument, so that plain
// tiles carry no division)
template <bool FIB>
__host__ __device__ inline int64_t first_off(const Strided& s, int i) {
    if (FIB) return (int64_t)(i % s.lo) * s.si + (int64_t)(i / s.lo) * s.hi;
    return (int64_t)i * s.si;
}

// One output tile of X = P Q (M x N, column-major with leading dimension
// ldx at x_off); q is read as (k, j).
struct GemmTile {
    Strided p, q;
    int64_t x_off;
    int M, N, K;
    int ti, tj;
    int ldx;
    int64_t xhi = 0;  // rows of X follow p.lo: i -> (i % lo) + (i / lo) xhi
};
template <bool FIB>
__host__ __device__ inline int64_t x_row(const GemmTile& g, int i) {
    if (FIB) return (int64_t)(i % g.p.lo) + (int64_t)(i / g.p.lo) * g.xhi;
    return (int64_t)i;
}

inline int tiles_of(int n) { return (n + kTile - 1) / kTile; }

typedef double v4d __attribute__((ext_vector_type(4)));

// The 16-row strip (rows i0 .. i0 + 15) of a 64 x 64 output tile of
// X = P Q, on the matrix cores (v_mfma_f64_16x16x4f64): acc[q] += the
// 16 x 16 block of columns j0 + 16 q, q < nq, scaled by `scale`.  Lane l
// supplies P(i0 + l % 16, k + l / 16) and Q(k + l / 16, j + l % 16) and holds
// rows l / 16 + 4 v, column l % 16 of each block.  Operands are read
// straight from memory with clamped indices (rows past M and columns past
// N repeat the last one and are never stored; k past K contributes zero),
// four k-steps' loads issued before their matrix-core steps.
template <bool FIB = false>
__device__ __forceinline__ void wave_strip(const double* __restrict__ base, const Strided& P,
                                           const Strided& Q, int M, int N, int K, int i0,
                                           int j0, int nq, double scale, v4d acc[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kk = lane >> 4;
    const int64_t prow = P.off + first_off<FIB>(P, min(i0 + r, M - 1));
    int64_t qcol[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) qcol[q] = Q.off + (int64_t)min(j0 + 16 * q + r, N - 1) * Q.sk;
    for (int k0 = 0; k0 < K; k0 += 16) {
        double a[4], b[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kc = k0 + 4 * u + kk;
            const int kl = kc < K ? kc : 0;
            a[u] = base[prow + (int64_t)kl * P.sk];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[u][q] = base[qcol[q] + (int64_t)kl * Q.si];
            a[u] = kc < K ? a[u] * scale : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k0 + 4 * u < K) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < nq) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][q], acc[q], 0, 0, 0);
            }
        }
    }
}

// Large tiles: one 16-deep K slab of a 64 x 64 tile product, P (64 x 16)
// and Q (16 x 64) staged through LDS by the whole workgroup (zero padded),
// acc[a][b] += for rows tx + 16 a, columns ty + 16 b (tx = t % 16,
// ty = t / 16) with vector FMAs.  The load pattern follows the unit-stride
// index so that neighbouring lanes read neighbouring addresses.
constexpr int kKStep = 16;
template <bool FIB = false>
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
        Ps[k][i] = (gi < M && gk < K) ? base[P.off + first_off<FIB>(P, gi) + (int64_t)gk * P.sk] : 0.0;
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

// A tile that fills most of its 64 x 64 footprint over a long K takes the
// LDS slabs (one load of each operand element serves the whole tile);
// others take the per-wave matrix-core strips.  Uniform per workgroup.
__device__ __forceinline__ bool tile_staged(int M, int N, int K, int i0, int j0) {
    return M - i0 >= 32 && N - j0 >= 32 && K >= 32;
}

// Column blocks of a 64 x 64 tile that lie inside an N-column matrix.
__device__ __forceinline__ int tile_nq(int N, int j0) { return min(4, (N - j0 + 15) >> 4); }

// X tiles = P Q, all operands in one arena (inputs and outputs of one
// launch must not overlap).  256 threads = four 16-row strips per tile,
// grid-stride over the tiles.
// Internal linkage: every translation unit that launches it has its copy.
// Stores a wave's strip (rows i0 .. i0 + 15) of a tile.
template <bool FIB>
__device__ __forceinline__ void strip_store(double* __restrict__ base, const GemmTile& g, int i0,
                                            int j0, int nq, const v4d acc[4]) {
    const int lane = threadIdx.x & 63;
    int64_t row[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) row[v] = x_row<FIB>(g, min(i0 + (lane >> 4) + 4 * v, g.M - 1));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + 16 * q + (lane & 15);
        if (q >= nq || j >= g.N) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + (lane >> 4) + 4 * v;
            if (i < g.M) base[g.x_off + row[v] + (int64_t)j * g.ldx] = acc[q][v];
        }
    }
}

// X tiles = P Q, all operands in one arena (inputs and outputs of one
// launch must not overlap).  256 threads; tiles are taken four at a time:
// when none of the four takes the LDS slabs, each wave computes one tile
// (all its strips), else the workgroup computes them one after another
// (four strips at once).  Internal linkage: every translation unit that
// launches it has its copy.
// One tile by one wave (all its strips).
template <bool FIB>
__device__ __forceinline__ void tile_by_wave(double* base, const GemmTile& g) {
    const int ti0 = g.ti * kTile, j0 = g.tj * kTile, nq = tile_nq(g.N, j0);
    for (int i0 = ti0; i0 < min(g.M, ti0 + kTile); i0 += 16) {
        v4d acc[4] = {};
        wave_strip<FIB>(base, g.p, g.q, g.M, g.N, g.K, i0, j0, nq, 1.0, acc);
        strip_store<FIB>(base, g, i0, j0, nq, acc);
    }
}

// One tile by the workgroup: LDS slabs for a full tile, else a strip per wave.
template <bool FIB>
__device__ __forceinline__ void tile_by_group(double* base, const GemmTile& g,
                                              double (*Ps)[kTile + 1], double (*Qs)[kTile + 1]) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ti0 = g.ti * kTile, j0 = g.tj * kTile;
    if (tile_staged(g.M, g.N, g.K, ti0, j0)) {
        double acc[4][4] = {};
        for (int k0 = 0; k0 < g.K; k0 += kKStep)
            tile_slab<FIB>(base, g.p, g.q, g.M, g.N, g.K, ti0, j0, k0, Ps, Qs, acc);
        const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = j0 + ty + 16 * b;
            if (j >= g.N) continue;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = ti0 + tx + 16 * a;
                if (i < g.M) base[g.x_off + x_row<FIB>(g, i) + (int64_t)j * g.ldx] = acc[a][b];
            }
        }
    } else {
        const int i0 = ti0 + 16 * w, nq = tile_nq(g.N, j0);
        if (i0 < g.M) {
            v4d acc[4] = {};
            wave_strip<FIB>(base, g.p, g.q, g.M, g.N, g.K, i0, j0, nq, 1.0, acc);
            strip_store<FIB>(base, g, i0, j0, nq, acc);
        }
    }
}

// X tiles = P Q, all operands in one arena (inputs and outputs of one
// launch must not overlap).  256 threads; tiles are taken four at a time:
// when none of the four takes the LDS slabs, each wave computes one tile
// (all its strips), else the workgroup computes them one after another
// (four strips at once).  Internal linkage: every translation unit that
// launches it has its copy.
static __global__ void __launch_bounds__(256)
    gemm_tiles_kernel(const GemmTile* __restrict__ tiles, int ntiles, double* base) {
    __shared__ double Ps[kKStep][kTile + 1];
    __shared__ double Qs[kKStep][kTile + 1];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int it = blockIdx.x * 4; it < ntiles; it += gridDim.x * 4) {
        int mine = 0;
        if (it + w < ntiles) {
            const GemmTile g = tiles[it + w];
            mine = tile_staged(g.M, g.N, g.K, g.ti * kTile, g.tj * kTile);
        }
        if (!__syncthreads_or(mine)) {  // one tile per wave
            if (it + w >= ntiles) continue;
            const GemmTile g = tiles[it + w];
            if (__builtin_amdgcn_readfirstlane(g.p.lo)) tile_by_wave<true>(base, g);
            else tile_by_wave<false>(base, g);
            continue;
        }
        for (int t = it; t < min(it + 4, ntiles); ++t) {
            const GemmTile g = tiles[t];
            if (__builtin_amdgcn_readfirstlane(g.p.lo)) tile_by_group<true>(base, g, Ps, Qs);
            else tile_by_group<false>(base, g, Ps, Qs);
        }
    }
}

}  // namespace tile
}  // namespace t3
