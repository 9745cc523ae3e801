// Device effective Hamiltonian: work-list construction and kernels.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "heff_dev.hpp"
#include "resident_ops.hpp"
#include "../host/host_threads.hpp"

namespace t3 {

namespace {
constexpr int kThreads = 512;  // 8 waves
constexpr int kDiagThreads = 256;
// LDS pool of a workgroup (elements), laid out per pair: the input block,
// then the first-stage outputs of a chunk, then its second-stage outputs
// (chunks are cut to fit); the stage-C partial tiles (256 values per wave)
// have their own region.  Three shapes of the persistent grid: WPC
// workgroups of 8 / WPC waves per CU (WPC = 1, 2, 4), each with the pool
// that lets WPC of them share a CU's 160 KiB of LDS (with the chunk
// records and partial tiles); smaller workgroups wait on memory and
// barriers independently of each other.
template <typename T, int WPC>
struct Grid {
    static constexpr int NT = kThreads / WPC;
    static constexpr int POOL = sizeof(T) == 8 ? (WPC == 4 ? 3072 : WPC == 2 ? 7680 : 16384)
                                               : (WPC == 4 ? 6912 : WPC == 2 ? 15360 : 32768);
};
constexpr int kShapes = 3;  // grid shapes: WPC = 1 << shape
// Terms of a chunk: one chunk record per thread, and one more group bound
// than groups.
constexpr int max_chunk_terms(int nt) { return nt - 1 < 128 ? nt - 1 : 128; }
constexpr int kPool64 = Grid<double, 1>::POOL, kPool32 = Grid<float, 1>::POOL;
constexpr int kMaxChunkTerms = 128;  // chunk records staged by one thread each
// Global scratch per workgroup for pairs whose stage outputs exceed LDS.
constexpr int64_t kGlobalChunkElems = 1 << 18;
const int hOrd[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// 16x16x4 matrix-core step, D = A B + C, one wave: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[row(l, v)][l % 16],
// v = 0..3, with row = 4 (l / 16) + v for FP32 and l / 16 + 4 v for FP64.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V run(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V run(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// FP32 raises the wave's issue priority over its runs of matrix-core steps
// (stage C: 242 against 247 ms per matvec on the QC-24 D = 512 step);
// FP64 does not (358 against 356).
template <typename T>
constexpr bool kMfmaPrio = sizeof(T) == 4;

// Operator values as the arithmetic type: FP64 and FP32 arenas as they
// are, BF16 arenas (the upper 16 bits of an FP32 value, per-block scaled)
// widened to FP32.
__device__ __forceinline__ double cvt(double v) { return v; }
__device__ __forceinline__ float cvt(float v) { return v; }
__device__ __forceinline__ float cvt(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }

// Geometry of a mode-p product of a three-leg block S (memory dims s0, s1,
// s2; first fastest) with a column-major (dn x s_p) operator block O:
//     Z[f, k'] = sum_k S[f, k] O[k', k],   f = (i_lo, i_hi) a fiber,
// with S[f, k] at i_lo + k lo + i_hi lo K and Z[f, k'] at i_lo + k' lo +
// i_hi lo dn.  Tiles are 16 fibers x 16 columns.
struct Geo {
    int lo, K, F, dn, tf, tiles, ksteps;
    double inv_lo;  // 1 / lo, for fiber_hi
};
// Values every lane of a wave agrees on, kept in scalar registers so that
// loop control stays scalar.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ Geo make_geo(int s0, int s1, int s2, int p, int dn) {
    s0 = uni(s0);
    s1 = uni(s1);
    s2 = uni(s2);
    p = uni(p);
    dn = uni(dn);
    Geo g;
    g.lo = p == 0 ? 1 : (p == 1 ? s0 : s0 * s1);
    g.K = p == 0 ? s0 : (p == 1 ? s1 : s2);
    const int hi = p == 0 ? s1 * s2 : (p == 1 ? s2 : 1);
    g.F = g.lo * hi;
    g.dn = dn;
    g.tf = (g.F + 15) >> 4;
    g.tiles = g.tf * ((dn + 15) >> 4);
    g.ksteps = (g.K + 3) >> 2;
    g.inv_lo = 1.0 / g.lo;
    return g;
}

// f / lo for 0 <= f < 2^31: a double-precision estimate (off by at most
// one) corrected by its remainder, a handful of vector instructions where
// an integer division takes several dozen.
__device__ __forceinline__ int fiber_hi(const Geo& g, int f) {
    int q = (int)((double)f * g.inv_lo);
    const int r = f - q * g.lo;
    q += (r >= g.lo) - (r < 0);
    return q;
}
// Offset of fiber f in a block whose mode-p leg has dimension n:
// i_lo + i_hi lo n = f + i_hi lo (n - 1).
__device__ __forceinline__ int fiber_off(const Geo& g, int f, int n) {
    return f + fiber_hi(g, f) * g.lo * (n - 1);
}

// Register blocking: a wave task covers one 16-fiber row tile and up to
// kTN column tiles of the output, so one A-operand load feeds kTN
// matrix-core steps.
constexpr int kTN = 4;

// Per-lane offsets of one row tile, computed once and reused for every
// operator and term that goes through it.  Lanes past the last fiber or
// column read a valid element (clamped index): their D rows and columns are
// never stored, so only the K tail needs masking.
struct RowLane {
    int soff;       // A operand offset at this lane's first k (clamped fiber)
    int ooff;       // B operand offset of column tile tn0 at this lane's first k
    int rowoff[4];  // output offset of the lane's D rows without the column term (-1: none)
    int tn0, ntn;   // column tiles [tn0, tn0 + ntn) of this task
};
template <typename T>
__device__ __forceinline__ RowLane row_lane(const Geo& g, int tf, int tn0) {
    RowLane L;
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int f = min(tf * 16 + i, g.F - 1);
    L.soff = fiber_off(g, f, g.K) + kk * g.lo;
    L.tn0 = tn0;
    const int ntn_all = (g.dn + 15) >> 4;
    L.ntn = min(kTN, ntn_all - tn0);
    const int n = min(tn0 * 16 + i, g.dn - 1);
    L.ooff = n + kk * g.dn;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int fr = tf * 16 + Mfma<T>::row(lane, v);
        L.rowoff[v] = fr < g.F ? fiber_off(g, fr, g.dn) : -1;
    }
    return L;
}

// One slab of at most NU k-steps (ns of them live, ns >= 1) of row_mac_n:
// loads first, then the matrix-core steps.
template <typename T, typename OT, int NTN, int NU, bool SCALE>
__device__ __forceinline__ void row_slab(const T* sp, const OT* op, const int (&coff)[NTN], int sst,
                                         int ost, int ns, T scale,
                                         typename Mfma<T>::V (&acc)[kTN]) {
    T av[NU];
    T bv[NU][NTN];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int uu = min(u, ns - 1);
        av[u] = sp[uu * sst];
#pragma unroll
        for (int j = 0; j < NTN; ++j) bv[u][j] = cvt(op[uu * ost + coff[j]]);
    }
    if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        if (u < ns) {
            T a = av[u];
            if (SCALE) a *= scale;
#pragma unroll
            for (int j = 0; j < NTN; ++j) acc[j] = Mfma<T>::run(a, bv[u][j], acc[j]);
        }
    }
    if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(0);
}

// acc[j] += (scale S)[row tile, :] O[column tile tn0 + j, :]^T over the whole K,
// j < NTN.  Loads are unconditional (clamped indices stay inside the
// blocks) so the compiler can keep several in flight; only the K tail is
// masked.
template <typename T, typename OT, int NTN, bool SCALE>
__device__ __forceinline__ void row_mac_n(const T* __restrict__ S, const OT* __restrict__ O,
                                          const Geo& g, const RowLane& L, T scale,
                                          typename Mfma<T>::V (&acc)[kTN]) {
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const T* sp = S + L.soff;
    const OT* op = O + L.ooff;
    const int sst = 4 * g.lo, ost = 4 * g.dn;
    const int n0 = min(L.tn0 * 16 + i, g.dn - 1);
    int coff[NTN];
#pragma unroll
    for (int j = 0; j < NTN; ++j) coff[j] = min((L.tn0 + j) * 16 + i, g.dn - 1) - n0;
    const int kfull = uni(g.K >> 2);
    // Full k-steps in slabs of KS: all loads of a slab are issued before its
    // matrix-core steps so their latencies overlap.  The last, short slab
    // loads unconditionally (re-reading its last step): guarded loads make
    // the compiler drain the memory counter before each step.
    constexpr int KS = 4;
    int s0 = 0;
    for (; s0 + KS <= kfull; s0 += KS) {
        row_slab<T, OT, NTN, KS, SCALE>(sp, op, coff, sst, ost, KS, scale, acc);
        sp += KS * sst;
        op += KS * ost;
    }
    if (s0 < kfull) {
        const int ns = uni(kfull - s0);
        row_slab<T, OT, NTN, KS - 1, SCALE>(sp, op, coff, sst, ost, ns, scale, acc);
        sp += ns * sst;
        op += ns * ost;
    }
    if (g.K & 3) {  // K tail: lanes past K contribute zero (their loads are skipped)
        const bool kv = (kfull << 2) + kk < g.K;
        const T* spc = kv ? sp : S + L.soff;
        const OT* opc = kv ? op : O + (L.ooff - kk * g.dn);  // row 0 stays inside the block
        T av = kv ? spc[0] : (T)0;
        if (SCALE) av *= scale;
#pragma unroll
        for (int j = 0; j < NTN; ++j) {
            const T bv = cvt(opc[coff[j]]);
            acc[j] = Mfma<T>::run(av, bv, acc[j]);
        }
    }
}

template <typename T, typename OT, bool SCALE>
__device__ __forceinline__ void row_mac(const T* __restrict__ S, const OT* __restrict__ O,
                                        const Geo& g, const RowLane& L, T scale,
                                        typename Mfma<T>::V (&acc)[kTN]) {
    switch (__builtin_amdgcn_readfirstlane(L.ntn)) {
        case 1: row_mac_n<T, OT, 1, SCALE>(S, O, g, L, scale, acc); break;
        case 2: row_mac_n<T, OT, 2, SCALE>(S, O, g, L, scale, acc); break;
        case 3: row_mac_n<T, OT, 3, SCALE>(S, O, g, L, scale, acc); break;
        default: row_mac_n<T, OT, 4, SCALE>(S, O, g, L, scale, acc); break;
    }
}

template <typename T>
__device__ __forceinline__ void row_put(T* __restrict__ Z, const Geo& g, const RowLane& L,
                                        const typename Mfma<T>::V (&acc)[kTN], bool add) {
    const int i = threadIdx.x & 15;
    int off[kTN][4];  // -1: no element
#pragma unroll
    for (int j = 0; j < kTN; ++j) {
        const int n = (L.tn0 + j) * 16 + i;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            off[j][v] = j < L.ntn && n < g.dn && L.rowoff[v] >= 0 ? L.rowoff[v] + n * g.lo : -1;
    }
    // all reads of an accumulating put are issued before its first write
    // (element 0 stands in for missing ones), so they overlap
    T old[kTN][4];
#pragma unroll
    for (int j = 0; j < kTN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) old[j][v] = add ? Z[max(off[j][v], 0)] : (T)0;
#pragma unroll
    for (int j = 0; j < kTN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (off[j][v] >= 0) Z[off[j][v]] = add ? old[j][v] + acc[j][v] : acc[j][v];
}

// Row blocking, the transpose of the scheme above: a wave task covers up to
// kTF row tiles and one column tile, so one B-operand (operator) load feeds
// kTF matrix-core steps.  Chosen when a geometry has many fibers and few
// output columns, where column blocking would reload every operator column
// tile once per row tile.
constexpr int kTF = 4;
struct RowBlk {
    int soff[kTF];   // A operand offsets (clamped fibers) at this lane's first k
    int ooff;        // B operand offset of the column tile at this lane's first k
    int tf0, ntf, tn;
};
template <typename T>
__device__ __forceinline__ RowBlk row_blk(const Geo& g, int tf0, int tn) {
    RowBlk B;
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    B.tf0 = tf0;
    B.ntf = min(kTF, g.tf - tf0);
    B.tn = tn;
#pragma unroll
    for (int r = 0; r < kTF; ++r) {
        const int f = min((tf0 + r) * 16 + i, g.F - 1);
        B.soff[r] = fiber_off(g, f, g.K) + kk * g.lo;
    }
    const int n = min(tn * 16 + i, g.dn - 1);
    B.ooff = n + kk * g.dn;
    return B;
}

// One slab of at most NU k-steps (ns of them live, ns >= 1) of blk_mac_n.
template <typename T, typename OT, int NTF, int NU, bool SCALE>
__device__ __forceinline__ void blk_slab(const T* S, const OT* op, const RowBlk& B, int so, int sst,
                                         int ost, int ns, T scale,
                                         typename Mfma<T>::V (&acc)[kTF]) {
    T bv[NU];
    T av[NU][NTF];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int uu = min(u, ns - 1);
        bv[u] = cvt(op[uu * ost]);
#pragma unroll
        for (int r = 0; r < NTF; ++r) av[u][r] = S[B.soff[r] + so + uu * sst];
    }
    if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        if (u < ns) {
#pragma unroll
            for (int r = 0; r < NTF; ++r) {
                T a = av[u][r];
                if (SCALE) a *= scale;
                acc[r] = Mfma<T>::run(a, bv[u], acc[r]);
            }
        }
    }
    if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(0);
}

// acc[r] += (scale S)[row tile tf0 + r, :] O[column tile, :]^T, r < NTF.
template <typename T, typename OT, int NTF, bool SCALE>
__device__ __forceinline__ void blk_mac_n(const T* __restrict__ S, const OT* __restrict__ O,
                                          const Geo& g, const RowBlk& B, T scale,
                                          typename Mfma<T>::V (&acc)[kTF]) {
    const int kk = (threadIdx.x & 63) >> 4;
    const int sst = 4 * g.lo, ost = 4 * g.dn;
    const OT* op = O + B.ooff;
    int so = 0;
    const int kfull = uni(g.K >> 2);
    constexpr int KS = 4;
    int s0 = 0;
    for (; s0 + KS <= kfull; s0 += KS) {
        blk_slab<T, OT, NTF, KS, SCALE>(S, op, B, so, sst, ost, KS, scale, acc);
        so += KS * sst;
        op += KS * ost;
    }
    if (s0 < kfull) {
        const int ns = uni(kfull - s0);
        blk_slab<T, OT, NTF, KS - 1, SCALE>(S, op, B, so, sst, ost, ns, scale, acc);
        so += ns * sst;
        op += ns * ost;
    }
    if (g.K & 3) {  // K tail: lanes past K contribute zero
        const bool kv = (kfull << 2) + kk < g.K;
        const T bv = cvt(kv ? op[0] : O[B.ooff - kk * g.dn]);  // row 0 stays inside the block
#pragma unroll
        for (int r = 0; r < NTF; ++r) {
            T a = kv ? S[B.soff[r] + so] : (T)0;
            if (SCALE) a *= scale;
            acc[r] = Mfma<T>::run(a, bv, acc[r]);
        }
    }
}

template <typename T, typename OT, bool SCALE>
__device__ __forceinline__ void blk_mac(const T* __restrict__ S, const OT* __restrict__ O,
                                        const Geo& g, const RowBlk& B, T scale,
                                        typename Mfma<T>::V (&acc)[kTF]) {
    switch (__builtin_amdgcn_readfirstlane(B.ntf)) {
        case 1: blk_mac_n<T, OT, 1, SCALE>(S, O, g, B, scale, acc); break;
        case 2: blk_mac_n<T, OT, 2, SCALE>(S, O, g, B, scale, acc); break;
        case 3: blk_mac_n<T, OT, 3, SCALE>(S, O, g, B, scale, acc); break;
        default: blk_mac_n<T, OT, 4, SCALE>(S, O, g, B, scale, acc); break;
    }
}

template <typename T>
__device__ __forceinline__ void blk_put(T* __restrict__ Z, const Geo& g, const RowBlk& B,
                                        const typename Mfma<T>::V (&acc)[kTF], bool add) {
    const int lane = threadIdx.x & 63;
    const int n = B.tn * 16 + (lane & 15);
    int off[kTF][4];  // -1: no element
#pragma unroll
    for (int r = 0; r < kTF; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int fr = (B.tf0 + r) * 16 + Mfma<T>::row(lane, v);
            off[r][v] = r < B.ntf && fr < g.F && n < g.dn ? fiber_off(g, fr, g.dn) + n * g.lo : -1;
        }
    // reads first, as in row_put
    T old[kTF][4];
#pragma unroll
    for (int r = 0; r < kTF; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) old[r][v] = add ? Z[max(off[r][v], 0)] : (T)0;
#pragma unroll
    for (int r = 0; r < kTF; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (off[r][v] >= 0) Z[off[r][v]] = add ? old[r][v] + acc[r][v] : acc[r][v];
}

// Stage C sums every term of a chunk into one set of accumulators.  It
// walks k-steps over groups of up to kTU terms, so that the operand loads of
// all terms of a group are in flight together: a per-term walk over a short
// K (legs of a few states) waits out one memory latency per term.
constexpr int kTU = 4;
__device__ __forceinline__ int64_t uni64(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename T, typename OT>
struct Terms {
    const T* s[kTU];   // stage inputs
    const OT* o[kTU];  // operator blocks
    T scale[kTU];
    int n;  // live terms, >= 1
};

template <typename T, typename OT, int NTN, bool SCALE>
__device__ __forceinline__ void row_mac_terms_n(const Terms<T, OT>& R, const Geo& g, const RowLane& L,
                                                typename Mfma<T>::V (&acc)[kTN]) {
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int n0 = min(L.tn0 * 16 + i, g.dn - 1);
    int coff[NTN];
#pragma unroll
    for (int j = 0; j < NTN; ++j) coff[j] = min((L.tn0 + j) * 16 + i, g.dn - 1) - n0;
    const int nks = uni(g.ksteps), nu = uni(R.n);
    for (int ks = 0; ks < nks; ++ks) {
        // lanes past K read k = 0 (in bounds) and contribute zero
        const bool kv = ks * 4 + kk < g.K;
        const int ao = L.soff + (kv ? ks * 4 * g.lo : -kk * g.lo);
        const int bo = L.ooff + (kv ? ks * 4 * g.dn : -kk * g.dn);
        T av[kTU];
        T bv[kTU][NTN];
#pragma unroll
        for (int u = 0; u < kTU; ++u) {
            av[u] = R.s[u][ao];
#pragma unroll
            for (int j = 0; j < NTN; ++j) bv[u][j] = cvt(R.o[u][bo + coff[j]]);
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int u = 0; u < kTU; ++u) {
            if (u < nu) {
                T a = kv ? av[u] : (T)0;
                if (SCALE) a *= R.scale[u];
#pragma unroll
                for (int j = 0; j < NTN; ++j) acc[j] = Mfma<T>::run(a, bv[u][j], acc[j]);
            }
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(0);
    }
}

template <typename T, typename OT, bool SCALE>
__device__ __forceinline__ void row_mac_terms(const Terms<T, OT>& R, const Geo& g, const RowLane& L,
                                              typename Mfma<T>::V (&acc)[kTN]) {
    switch (__builtin_amdgcn_readfirstlane(L.ntn)) {
        case 1: row_mac_terms_n<T, OT, 1, SCALE>(R, g, L, acc); break;
        case 2: row_mac_terms_n<T, OT, 2, SCALE>(R, g, L, acc); break;
        case 3: row_mac_terms_n<T, OT, 3, SCALE>(R, g, L, acc); break;
        default: row_mac_terms_n<T, OT, 4, SCALE>(R, g, L, acc); break;
    }
}

// Fused middle stage.  When the middle leg of a three-stage order has at
// most kFuseN states in and out, its product is not materialised: stage C
// builds each A operand on the fly from the first-stage output,
//     W2[f, k] = sum_b Bs[b'(f), b] W1[f(b), k],
// with Bs = alpha_t O_b (at most kFuseN x kFuseN values per term, staged
// in LDS per chunk, padded to kFuseN x kFuseN), so the middle stage has no
// barrier, no vector-unit tasks and no second-stage outputs in LDS (a chunk
// then holds many more terms).
constexpr int kFuseN = 4;
constexpr int kFuseStage = kFuseN * kFuseN;  // staged values per term
constexpr int kFTU = 2;                      // terms per k-step walk (register budget)
struct FLane {
    int w1off;  // first-stage output offset of this lane's fiber at b = 0, k = its first k
    int bpr;    // the fiber's index b' on the middle leg
};
// Fiber f of the stage C geometry over memory dims (e0, e1, e2) (contracted
// position pc, middle leg at position pb) mapped to the first-stage output,
// whose memory dims equal these except at pb (f0, f1 are its first two).
__device__ __forceinline__ FLane fused_lane(const Geo& g, int tf, int e0, int e1, int f0, int f1,
                                            int pb, int pc) {
    FLane L;
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int f = min(tf * 16 + i, g.F - 1);
    int c0, c1, c2;
    if (pc == 0) {
        c0 = 0;
        c1 = f % e1;
        c2 = f / e1;
    } else if (pc == 1) {
        c0 = f % e0;
        c1 = 0;
        c2 = f / e0;
    } else {
        c0 = f % e0;
        c1 = f / e0;
        c2 = 0;
    }
    L.bpr = pb == 0 ? c0 : (pb == 1 ? c1 : c2);
    c0 = pb == 0 ? 0 : c0;
    c1 = pb == 1 ? 0 : c1;
    c2 = pb == 2 ? 0 : c2;
    const int sk = pc == 0 ? 1 : (pc == 1 ? f0 : f0 * f1);
    L.w1off = c0 + f0 * (c1 + f1 * c2) + kk * sk;
    return L;
}
// Stage C items of the fused form: groups of terms sharing their last
// operator (the A operand of a group sums its terms' W2).
template <typename T, typename OT>
struct FTerms {
    int tb[kFTU], te[kFTU];  // terms of each group (chunk-relative)
    const OT* o[kFTU];       // stage C operator blocks
    int n;                   // live groups, >= 1
};
// acc[j] += sum over the groups g of R and their terms t of
// W2_t[row tile, :] O_c,g[column tile tn0 + j, :]^T, W2_t built from W1 as
// above (W1 + slot_t w1; bst: the staged blocks; Kb: middle-leg states in,
// sb: its stride in W1, sk4: four k-steps of the contracted leg in W1).
template <typename T, typename OT, int NTN>
__device__ __forceinline__ void row_mac_fused_n(const FTerms<T, OT>& R, const Geo& g,
                                                const RowLane& L, const FLane& F, const T* W1,
                                                int w1, const T* bst, const int* slot, int Kb,
                                                int sb, int sk4, typename Mfma<T>::V (&acc)[kTN]) {
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int n0 = min(L.tn0 * 16 + i, g.dn - 1);
    int coff[NTN];
#pragma unroll
    for (int j = 0; j < NTN; ++j) coff[j] = min((L.tn0 + j) * 16 + i, g.dn - 1) - n0;
    const int nks = uni(g.ksteps), nu = uni(R.n);
    const int sk = sk4 >> 2;
    for (int ks = 0; ks < nks; ++ks) {
        // lanes past K read k = 0 (in bounds) and contribute zero
        const bool kv = ks * 4 + kk < g.K;
        const int ao = F.w1off + (kv ? ks * sk4 : -kk * sk);
        const int bo = L.ooff + (kv ? ks * 4 * g.dn : -kk * g.dn);
        T av[kFTU];
        T bv[kFTU][NTN];
        // operator loads first: their memory latency overlaps the LDS work
        // of the A operands
#pragma unroll
        for (int u = 0; u < kFTU; ++u)
#pragma unroll
            for (int j = 0; j < NTN; ++j) bv[u][j] = cvt(R.o[u][bo + coff[j]]);
#pragma unroll
        for (int u = 0; u < kFTU; ++u) {
            T a = 0;
            for (int t = R.tb[u]; t < R.te[u]; ++t) {
                const T* sp = W1 + uni(slot[t]) * w1 + ao;
                const T* bp = bst + t * kFuseStage + F.bpr;
#pragma unroll
                for (int q = 0; q < kFuseN; ++q)
                    if (q < Kb) a = fma(bp[kFuseN * q], sp[q * sb], a);
            }
            av[u] = a;
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int u = 0; u < kFTU; ++u) {
            if (u < nu) {
                const T a = kv ? av[u] : (T)0;
#pragma unroll
                for (int j = 0; j < NTN; ++j) acc[j] = Mfma<T>::run(a, bv[u][j], acc[j]);
            }
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(0);
    }
}

template <typename T, typename OT>
__device__ __forceinline__ void row_mac_fused(const FTerms<T, OT>& R, const Geo& g, const RowLane& L,
                                              const FLane& F, const T* W1, int w1, const T* bst,
                                              const int* slot, int Kb, int sb, int sk4,
                                              typename Mfma<T>::V (&acc)[kTN]) {
    switch (__builtin_amdgcn_readfirstlane(L.ntn)) {
        case 1: row_mac_fused_n<T, OT, 1>(R, g, L, F, W1, w1, bst, slot, Kb, sb, sk4, acc); break;
        case 2: row_mac_fused_n<T, OT, 2>(R, g, L, F, W1, w1, bst, slot, Kb, sb, sk4, acc); break;
        case 3: row_mac_fused_n<T, OT, 3>(R, g, L, F, W1, w1, bst, slot, Kb, sb, sk4, acc); break;
        default: row_mac_fused_n<T, OT, 4>(R, g, L, F, W1, w1, bst, slot, Kb, sb, sk4, acc); break;
    }
}

template <typename T, typename OT, int NTF, bool SCALE>
__device__ __forceinline__ void blk_mac_terms_n(const Terms<T, OT>& R, const Geo& g, const RowBlk& B,
                                                typename Mfma<T>::V (&acc)[kTF]) {
    const int kk = (threadIdx.x & 63) >> 4;
    const int nks = uni(g.ksteps), nu = uni(R.n);
    for (int ks = 0; ks < nks; ++ks) {
        const bool kv = ks * 4 + kk < g.K;
        const int so = kv ? ks * 4 * g.lo : -kk * g.lo;
        const int bo = B.ooff + (kv ? ks * 4 * g.dn : -kk * g.dn);
        T bv[kTU];
        T av[kTU][NTF];
#pragma unroll
        for (int u = 0; u < kTU; ++u) {
            bv[u] = cvt(R.o[u][bo]);
#pragma unroll
            for (int r = 0; r < NTF; ++r) av[u][r] = R.s[u][B.soff[r] + so];
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int u = 0; u < kTU; ++u) {
            if (u < nu) {
#pragma unroll
                for (int r = 0; r < NTF; ++r) {
                    T a = kv ? av[u][r] : (T)0;
                    if (SCALE) a *= R.scale[u];
                    acc[r] = Mfma<T>::run(a, bv[u], acc[r]);
                }
            }
        }
        if constexpr (kMfmaPrio<T>) __builtin_amdgcn_s_setprio(0);
    }
}

template <typename T, typename OT, bool SCALE>
__device__ __forceinline__ void blk_mac_terms(const Terms<T, OT>& R, const Geo& g, const RowBlk& B,
                                              typename Mfma<T>::V (&acc)[kTF]) {
    switch (__builtin_amdgcn_readfirstlane(B.ntf)) {
        case 1: blk_mac_terms_n<T, OT, 1, SCALE>(R, g, B, acc); break;
        case 2: blk_mac_terms_n<T, OT, 2, SCALE>(R, g, B, acc); break;
        case 3: blk_mac_terms_n<T, OT, 3, SCALE>(R, g, B, acc); break;
        default: blk_mac_terms_n<T, OT, 4, SCALE>(R, g, B, acc); break;
    }
}

// Vector-unit path for stages whose output leg has at most kVN states (a
// matrix-core tile would be mostly padding there): one lane per fiber, the
// fiber's dn outputs in registers, the operator values shared by the wave
// (uniform loads).  A task is 64 fibers; items are split over groups of
// waves as in the matrix-core path.
constexpr int kVN = 4;
struct VLane {
    int soff;  // S offset of the lane's fiber (clamped to the last fiber)
    int zoff;  // Z offset of the lane's fiber
    int live;  // the fiber exists
};
__device__ __forceinline__ VLane vlane(const Geo& g, int ft) {
    VLane L;
    const int f = ft * 64 + (threadIdx.x & 63);
    const int fc = min(f, g.F - 1);
    const int hi = fiber_hi(g, fc);
    L.soff = fc + hi * g.lo * (g.K - 1);
    L.zoff = fc + hi * g.lo * (g.dn - 1);
    L.live = f < g.F;
    return L;
}
struct VTasks {
    int nft, G, total;  // fiber tiles, item groups, tasks
};
__device__ __forceinline__ VTasks vtasks(const Geo& g, int nw) {
    VTasks t;
    t.nft = (g.F + 63) >> 6;
    t.G = t.nft >= nw ? 1 : nw / t.nft;
    t.total = t.nft * t.G;
    return t;
}

// acc[n] += sum_k (scale S)[f, k] O[n, k], n < NB = dn (O is dn x K, column-major)
template <typename T, typename OT, int NB, bool SCALE>
__device__ __forceinline__ void vmac_n(const T* __restrict__ S, const OT* __restrict__ O,
                                       const Geo& g, const VLane& L, T scale,
                                       typename Mfma<T>::V& acc) {
    const T* sp = S + L.soff;
    const int K = uni(g.K), lo = uni(g.lo);
    int k = 0;
    for (; k + 4 <= K; k += 4) {
        T a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = sp[(k + u) * lo];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const T av = SCALE ? a[u] * scale : a[u];
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[n] = fma(av, (T)cvt(O[(k + u) * NB + n]), acc[n]);
        }
    }
    for (; k < K; ++k) {
        const T av = SCALE ? sp[k * lo] * scale : sp[k * lo];
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[n] = fma(av, (T)cvt(O[k * NB + n]), acc[n]);
    }
}

template <typename T, typename OT, bool SCALE>
__device__ __forceinline__ void vmac(const T* __restrict__ S, const OT* __restrict__ O,
                                     const Geo& g, const VLane& L, T scale,
                                     typename Mfma<T>::V& acc) {
    switch (uni(g.dn)) {
        case 1: vmac_n<T, OT, 1, SCALE>(S, O, g, L, scale, acc); break;
        case 2: vmac_n<T, OT, 2, SCALE>(S, O, g, L, scale, acc); break;
        case 3: vmac_n<T, OT, 3, SCALE>(S, O, g, L, scale, acc); break;
        default: vmac_n<T, OT, 4, SCALE>(S, O, g, L, scale, acc); break;
    }
}

// Z[f, n] (+)= acc[n] for the lane's fiber; an accumulating put reads
// before it writes.
template <typename T>
__device__ __forceinline__ void vput(T* __restrict__ Z, const Geo& g, const VLane& L,
                                     const typename Mfma<T>::V& acc, bool add) {
    if (!L.live) return;
    T* z = Z + L.zoff;
    const int lo = g.lo, dn = uni(g.dn);
    T old[kVN];
#pragma unroll
    for (int n = 0; n < kVN; ++n) old[n] = add && n < dn ? z[n * lo] : (T)0;
#pragma unroll
    for (int n = 0; n < kVN; ++n)
        if (n < dn) z[n * lo] = add ? old[n] + acc[n] : acc[n];
}

// Tasks of one stage: (row tile, column-tile group, item group).
struct StageTasks {
    int tf, tng, G;  // row tiles, column-tile groups, item groups
    int total;
};
__device__ __forceinline__ StageTasks stage_tasks(const Geo& g, int nw) {
    StageTasks t;
    t.tf = g.tf;
    t.tng = (((g.dn + 15) >> 4) + kTN - 1) / kTN;
    const int base = t.tf * t.tng;
    t.G = base >= nw ? 1 : nw / base;
    t.total = base * t.G;
    return t;
}
// Row-blocked tasks: (row-tile group, column tile, item group); tf counts
// row-tile groups here and tng column tiles.
__device__ __forceinline__ StageTasks stage_tasks_rb(const Geo& g, int nw) {
    StageTasks t;
    t.tf = (g.tf + kTF - 1) / kTF;
    t.tng = (g.dn + 15) >> 4;
    const int base = t.tf * t.tng;
    t.G = base >= nw ? 1 : nw / base;
    t.total = base * t.G;
    return t;
}
// Row blocking pays when the output has few column tiles and several row
// tiles (every operator tile is then loaded once per row-tile group).
__device__ __forceinline__ bool use_rows(const Geo& g) {
    return ((g.dn + 15) >> 4) <= 2 && g.tf >= 2;
}

// Operator offsets, prefactors, first-stage slots and term groups of one
// chunk, staged in LDS (double-buffered: the next chunk's are loaded while
// this one runs).
struct ChunkRec {
    int64_t opa[kMaxChunkTerms];  // per run
    int64_t opb[kMaxChunkTerms];  // per term
    int64_t opc[kMaxChunkTerms];  // per group
    double alpha[kMaxChunkTerms];  // per term
    int slot[kMaxChunkTerms];      // per term
    int cbeg[kMaxChunkTerms + 1];  // terms of group g: [cbeg[g], cbeg[g + 1])
};

// Stage B sums groups of terms only in the FP32 kernel (the host groups
// terms on that path only).
template <typename T>
constexpr bool kGroupB = sizeof(T) == 4;
// The fused middle stage is built where stage outputs live in LDS.
template <typename T>
constexpr bool kFused = true;

// Terms t0, t0 + G, ... (at most kTU of them) of a chunk as stage C inputs.
template <typename T, typename OT>
__device__ __forceinline__ Terms<T, OT> c_terms(const ChunkRec* rec, int t0, int G, int nt,
                                                int isdmrg, const T* W1, int w1, const T* W2,
                                                int w2, const OT* opc) {
    Terms<T, OT> R;
    R.n = uni(min(kTU, (nt - t0 + G - 1) / G));
#pragma unroll
    for (int u = 0; u < kTU; ++u) {
        const int t = t0 + min(u, R.n - 1) * G;
        R.o[u] = opc + uni64(rec->opc[t]);
        if (isdmrg) {
            R.s[u] = W1 + uni(rec->slot[t]) * w1;
            R.scale[u] = (T)rec->alpha[t];
        } else {
            R.s[u] = W2 + t * w2;
            R.scale[u] = (T)1;
        }
    }
    return R;
}

// All chunks of one pair.  Stage A: first-stage outputs of the chunk's
// distinct first operators; stage B: second-stage output of every term
// group; stage C: Y += sum over the groups of O_c x_c W2_g.  Each stage
// picks its engine from its geometry (vector unit for output legs of at
// most kVN states, else matrix cores, row-blocked for narrow outputs).
// Waves take (tile, item group) tasks: a wave keeps its tile's lane offsets
// and walks the items of its group; when a geometry has fewer tiles than
// waves, items are split over groups (stage C partial tiles are then
// reduced in a fixed order).
template <typename T, typename OT, int NT, bool FZ>
__device__ __forceinline__ void pair_chunks(const DevPair& P, const int* bd,
                                            const DevChunk* __restrict__ chunks,
                                            const int* __restrict__ runs,
                                            const DevTerm* __restrict__ terms,
                                            const OT* __restrict__ op0, const OT* __restrict__ op1,
                                            const OT* __restrict__ op2, const T* X, T* W1, T* Y,
                                            T* part, ChunkRec* recs, int isdmrg, int m0, int m1,
                                            int m2) {
    typedef typename Mfma<T>::V V;
    constexpr int NW = NT / 64;
    const int wave = uni(threadIdx.x >> 6);  // wave-uniform: task loops stay scalar
    const int lane = threadIdx.x & 63;
    const int a = uni(P.ord[0]), b = uni(P.ord[1]), c = uni(isdmrg ? P.ord[1] : P.ord[2]);
    // Leg dimensions through the stages.  Only constant indices into
    // arrays, selects otherwise: a runtime index would put the array in
    // scratch memory.
    const int pd0 = uni(P.d[0]), pd1 = uni(P.d[1]), pd2 = uni(P.d[2]);
    const int bd0 = uni(bd[0]), bd1 = uni(bd[1]), bd2 = uni(bd[2]);
    auto posof = [&](int k) { return k == m0 ? 0 : (k == m1 ? 1 : 2); };
    auto bdof = [&](int k) { return k == 0 ? bd0 : (k == 1 ? bd1 : bd2); };
    auto pdof = [&](int k) { return k == 0 ? pd0 : (k == 1 ? pd1 : pd2); };
    int cur0 = pdof(m0), cur1 = pdof(m1), cur2 = pdof(m2);
    auto setcur = [&](int p, int v) {
        cur0 = p == 0 ? v : cur0;
        cur1 = p == 1 ? v : cur1;
        cur2 = p == 2 ? v : cur2;
    };
    const Geo gA = make_geo(cur0, cur1, cur2, posof(a), bdof(a));
    setcur(posof(a), bdof(a));
    const int w1d0 = cur0, w1d1 = cur1;  // first-stage output dims (first two)
    Geo gB = gA;
    if (!isdmrg) {
        gB = make_geo(cur0, cur1, cur2, posof(b), bdof(b));
        setcur(posof(b), bdof(b));
    }
    const Geo gC = make_geo(cur0, cur1, cur2, posof(c), bdof(c));
    // fused middle stage (three-leg objects only): its input states, its
    // stride in W1 and four k-steps of stage C's contracted leg in W1
    const bool fz = FZ && uni(P.fuse) != 0;
    const int fKb = gB.K, fpb = posof(b), fpc = posof(c);
    const int fsb = fpb == 0 ? 1 : (fpb == 1 ? w1d0 : w1d0 * w1d1);
    const int fsk4 = 4 * (fpc == 0 ? 1 : (fpc == 1 ? w1d0 : w1d0 * w1d1));
    // stages whose output leg has at most kVN states take the vector-unit path
    const bool vA = gA.dn <= kVN, vB = gB.dn <= kVN, vC = gC.dn <= kVN;
    const VTasks tA = vtasks(gA, NW), tB = vtasks(gB, NW), tC = vtasks(gC, NW);
    const bool rbA = use_rows(gA), rbB = use_rows(gB);
    const StageTasks kA = rbA ? stage_tasks_rb(gA, NW) : stage_tasks(gA, NW);
    const StageTasks kB = rbB ? stage_tasks_rb(gB, NW) : stage_tasks(gB, NW);
    StageTasks kC = stage_tasks(gC, NW);
    // stage C goes row-blocked only when its tasks alone occupy every wave
    // (no item groups, hence no partial-tile reduction)
    const StageTasks kCr = stage_tasks_rb(gC, NW);
    const bool rbC = !fz && use_rows(gC) && kCr.tf * kCr.tng >= NW;
    // stage C item groups are reduced through LDS: keep the partial tiles
    // within NW x 256 values
    const int ntnC = (gC.dn + 15) >> 4;
    kC.G = gC.tf * ntnC >= NW || kC.tng > 1 ? 1 : NW / (gC.tf * ntnC);
    kC.total = kC.tf * kC.tng * kC.G;
    // On the vector-unit path, when every wave has at most one stage C task,
    // its accumulators stay in registers across the pair's chunks and Y is
    // read and written once per pair instead of once per chunk.  (Kept
    // across chunks on the matrix-core path too, they took 46 registers of
    // the kernel's 256 and saved nothing measurable.)
    const bool pCv = vC && tC.total <= NW;
    const VLane LV = vlane(gC, wave % tC.nft);
    const int grpV = wave / tC.nft;
    const int baseC = kC.tf * kC.tng;
    V accV = V{0, 0, 0, 0};
    // stage C partial tiles of item groups, summed in group order into Y
    auto reduce_parts = [&]() {
        __syncthreads();
        for (int r = wave; r < kC.tf; r += NW) {
            const RowLane L = row_lane<T>(gC, r, 0);
            V acc[kTN];
#pragma unroll
            for (int q = 0; q < kTN; ++q) {
                acc[q] = V{0, 0, 0, 0};
                if (q < L.ntn)
                    for (int grp = 0; grp < kC.G; ++grp) {
                        const T* pp = part + (((grp * kC.tf + r) * ntnC + q) * 64 + lane) * 4;
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[q][v] += pp[v];
                    }
            }
            row_put<T>(Y, gC, L, acc, true);
        }
    };
    auto store_part = [&](const V (&acc)[kTN], const RowLane& L, int grp, int r) {
#pragma unroll
        for (int q = 0; q < kTN; ++q) {
            if (q < L.ntn) {
                T* pp = part + (((grp * kC.tf + r) * ntnC + q) * 64 + lane) * 4;
#pragma unroll
                for (int v = 0; v < 4; ++v) pp[v] = acc[q][v];
            }
        }
    };
    // vector-unit stage C partials: (group, fiber tile) x 64 lanes x 4 values
    auto vstore_part = [&](const V& acc, int grp, int ft) {
        T* pp = part + ((grp * tC.nft + ft) * 64 + lane) * 4;
#pragma unroll
        for (int n = 0; n < 4; ++n) pp[n] = acc[n];
    };
    auto vreduce_parts = [&]() {
        __syncthreads();
        for (int ft = wave; ft < tC.nft; ft += NW) {
            const VLane L = vlane(gC, ft);
            V acc = V{0, 0, 0, 0};
            for (int grp = 0; grp < tC.G; ++grp) {
                const T* pp = part + ((grp * tC.nft + ft) * 64 + lane) * 4;
#pragma unroll
                for (int n = 0; n < 4; ++n) acc[n] += pp[n];
            }
            vput<T>(Y, gC, L, acc, true);
        }
    };
    const OT* opa = a == 0 ? op0 : (a == 1 ? op1 : op2);
    const OT* opbb = b == 0 ? op0 : (b == 1 ? op1 : op2);
    const OT* opc = c == 0 ? op0 : (c == 1 ? op1 : op2);
    const int cb = uni(P.chunk_begin), ce = uni(P.chunk_end);
    const int w1 = uni(P.w1), w2 = uni(P.w2);
    for (int ci = cb; ci < ce; ++ci) {
        const DevChunk C = chunks[ci];
        T* const W2 = W1 + uni(C.w2_off);  // this chunk's second-stage outputs
        const int nr = uni(C.run_end - C.run_begin);
        const int nt = uni(C.term_end - C.term_begin);
        const int ncs = uni(C.ncs);  // stage C items (groups of terms)
        // chunk records in LDS (double-buffered; the first chunk of a pair
        // loads its own, the others were loaded during the previous chunk)
        ChunkRec* const rec = recs + ((ci - cb) & 1);
        if (ci == cb) {
            if (threadIdx.x < nt) {
                const DevTerm& tm = terms[uni(C.term_begin) + threadIdx.x];
                rec->opb[threadIdx.x] = tm.op[b];
                rec->alpha[threadIdx.x] = tm.alpha;
                rec->slot[threadIdx.x] = tm.slot;
            }
            if (threadIdx.x <= ncs) {
                const int g0 = runs[uni(C.run_end) + threadIdx.x];
                rec->cbeg[threadIdx.x] = g0;
                if (threadIdx.x < ncs) rec->opc[threadIdx.x] = terms[uni(C.term_begin) + g0].op[c];
            }
            if (threadIdx.x < nr) rec->opa[threadIdx.x] = terms[runs[uni(C.run_begin) + threadIdx.x]].op[a];
            __syncthreads();
        }
        // fused middle stage: alpha_t O_b of every term, loaded now and
        // staged in LDS (after W1) at the end of stage A
        constexpr int kFuseRegs = kMaxChunkTerms * kFuseStage / NT;
        if (vA) {  // stage A, vector unit
            for (int task = wave; task < tA.total; task += NW) {
                const VLane L = vlane(gA, task % tA.nft);
                for (int j = task / tA.nft; j < nr; j += tA.G) {
                    V acc = V{0, 0, 0, 0};
                    vmac<T, OT, false>(X, opa + uni64(rec->opa[j]), gA, L, (T)1, acc);
                    vput<T>(W1 + j * w1, gA, L, acc, false);
                }
            }
        } else if (rbA) {  // stage A, row-blocked
            const int base = kA.tf * kA.tng;
            for (int task = wave; task < kA.total; task += NW) {
                const int r = task % base, grp = task / base;
                const RowBlk L = row_blk<T>(gA, (r % kA.tf) * kTF, r / kA.tf);
                for (int j = grp; j < nr; j += kA.G) {
                    V acc[kTF];
#pragma unroll
                    for (int q = 0; q < kTF; ++q) acc[q] = V{0, 0, 0, 0};
                    blk_mac<T, OT, false>(X, opa + rec->opa[j], gA, L, (T)1, acc);
                    blk_put<T>(W1 + j * w1, gA, L, acc, false);
                }
            }
        } else {  // stage A: W1[j] = O_a[j] x_a X
            const int base = kA.tf * kA.tng;
            for (int task = wave; task < kA.total; task += NW) {
                const int r = task % base, grp = task / base;
                const RowLane L = row_lane<T>(gA, r % kA.tf, (r / kA.tf) * kTN);
                for (int j = grp; j < nr; j += kA.G) {
                    V acc[kTN];
#pragma unroll
                    for (int q = 0; q < kTN; ++q) acc[q] = V{0, 0, 0, 0};
                    row_mac<T, OT, false>(X, opa + rec->opa[j], gA, L, (T)1, acc);
                    row_put<T>(W1 + j * w1, gA, L, acc, false);
                }
            }
        }
        if (fz) {
            const int dnb = uni(gB.dn);
#pragma unroll
            for (int r = 0; r < kFuseRegs; ++r) {
                const int idx = threadIdx.x + r * NT;
                const int t = idx / kFuseStage, e = idx % kFuseStage;
                const int bp = e % kFuseN, bq = e / kFuseN;
                T v = (T)0;
                if (t < nt && bp < dnb && bq < fKb)
                    v = (T)rec->alpha[t] * (T)cvt(opbb[rec->opb[t] + bp + bq * dnb]);
                if (idx < nt * kFuseStage) W2[idx] = v;
            }
        }
        __syncthreads();
        if (!isdmrg && !fz && vB) {  // stage B, vector unit
            for (int task = wave; task < tB.total; task += NW) {
                const VLane L = vlane(gB, task % tB.nft);
                for (int g = task / tB.nft; g < ncs; g += tB.G) {
                    V acc = V{0, 0, 0, 0};
                    if constexpr (kGroupB<T>) {
                        for (int t = uni(rec->cbeg[g]); t < uni(rec->cbeg[g + 1]); ++t)
                            vmac<T, OT, true>(W1 + uni(rec->slot[t]) * w1, opbb + uni64(rec->opb[t]), gB,
                                          L, (T)rec->alpha[t], acc);
                    } else {  // one term per group
                        vmac<T, OT, true>(W1 + uni(rec->slot[g]) * w1, opbb + uni64(rec->opb[g]), gB, L,
                                      (T)rec->alpha[g], acc);
                    }
                    vput<T>(W2 + g * w2, gB, L, acc, false);
                }
            }
            __syncthreads();
        } else if (!isdmrg && !fz && rbB) {  // stage B, row-blocked
            const int base = kB.tf * kB.tng;
            for (int task = wave; task < kB.total; task += NW) {
                const int r = task % base, grp = task / base;
                const RowBlk L = row_blk<T>(gB, (r % kB.tf) * kTF, r / kB.tf);
                for (int g = grp; g < ncs; g += kB.G) {
                    V acc[kTF];
#pragma unroll
                    for (int q = 0; q < kTF; ++q) acc[q] = V{0, 0, 0, 0};
                    if constexpr (kGroupB<T>) {
                        for (int t = uni(rec->cbeg[g]); t < uni(rec->cbeg[g + 1]); ++t)
                            blk_mac<T, OT, true>(W1 + rec->slot[t] * w1, opbb + rec->opb[t], gB, L,
                                             (T)rec->alpha[t], acc);
                    } else {  // one term per group
                        blk_mac<T, OT, true>(W1 + rec->slot[g] * w1, opbb + rec->opb[g], gB, L,
                                         (T)rec->alpha[g], acc);
                    }
                    blk_put<T>(W2 + g * w2, gB, L, acc, false);
                }
            }
            __syncthreads();
        } else if (!isdmrg && !fz) {  // stage B: W2[t] = alpha_t O_b[t] x_b W1[slot t]
            const int base = kB.tf * kB.tng;
            for (int task = wave; task < kB.total; task += NW) {
                const int r = task % base, grp = task / base;
                const RowLane L = row_lane<T>(gB, r % kB.tf, (r / kB.tf) * kTN);
                for (int g = grp; g < ncs; g += kB.G) {
                    V acc[kTN];
#pragma unroll
                    for (int q = 0; q < kTN; ++q) acc[q] = V{0, 0, 0, 0};
                    if constexpr (kGroupB<T>) {
                        for (int t = uni(rec->cbeg[g]); t < uni(rec->cbeg[g + 1]); ++t)
                            row_mac<T, OT, true>(W1 + rec->slot[t] * w1, opbb + rec->opb[t], gB, L,
                                             (T)rec->alpha[t], acc);
                    } else {  // one term per group
                        row_mac<T, OT, true>(W1 + rec->slot[g] * w1, opbb + rec->opb[g], gB, L,
                                         (T)rec->alpha[g], acc);
                    }
                    row_put<T>(W2 + g * w2, gB, L, acc, false);
                }
            }
            __syncthreads();
        }
        // The next chunk's records: loaded now, stored into the other
        // buffer before the barrier that ends stage C (their latency
        // overlaps stage C).
        const bool pf = ci + 1 < ce;
        int64_t pfb = 0, pfc = 0, pfa = 0;
        double pfal = 0;
        int pfsl = 0, pfnt = 0, pfnr = 0, pfnc = 0, pfg = 0;
        if (pf) {
            const DevChunk C2 = chunks[ci + 1];
            pfnt = uni(C2.term_end - C2.term_begin);
            pfnr = uni(C2.run_end - C2.run_begin);
            pfnc = uni(C2.ncs);
            if (threadIdx.x < pfnt) {
                const DevTerm& tm = terms[uni(C2.term_begin) + threadIdx.x];
                pfb = tm.op[b];
                pfal = tm.alpha;
                pfsl = tm.slot;
            }
            if (threadIdx.x <= pfnc) {
                pfg = runs[uni(C2.run_end) + threadIdx.x];
                if (threadIdx.x < pfnc) pfc = terms[uni(C2.term_begin) + pfg].op[c];
            }
            if (threadIdx.x < pfnr) pfa = terms[runs[uni(C2.run_begin) + threadIdx.x]].op[a];
        }
        if (fz) {  // stage C with the middle stage fused, over groups of terms
            auto fterms = [&](int g0, int G) {
                FTerms<T, OT> R;
                R.n = uni(min(kFTU, (ncs - g0 + G - 1) / G));
#pragma unroll
                for (int u = 0; u < kFTU; ++u) {
                    const int gg = g0 + min(u, R.n - 1) * G;
                    R.tb[u] = uni(rec->cbeg[gg]);
                    R.te[u] = uni(rec->cbeg[gg + 1]);
                    R.o[u] = opc + uni64(rec->opc[gg]);
                }
                return R;
            };
            {  // Y is read and written once per chunk (a fused pair has few chunks)
                for (int task = wave; task < kC.total; task += NW) {
                    const int r = task % baseC, grp = task / baseC;
                    const RowLane L = row_lane<T>(gC, r % kC.tf, (r / kC.tf) * kTN);
                    const FLane FL = fused_lane(gC, r % kC.tf, cur0, cur1, w1d0, w1d1, fpb, fpc);
                    V acc[kTN];
#pragma unroll
                    for (int q = 0; q < kTN; ++q) acc[q] = V{0, 0, 0, 0};
                    for (int g0 = grp; g0 < ncs; g0 += kC.G * kFTU)
                        row_mac_fused<T, OT>(fterms(g0, kC.G), gC, L, FL, W1, w1, W2, rec->slot,
                                             fKb, fsb, fsk4, acc);
                    if (kC.G == 1) row_put<T>(Y, gC, L, acc, true);
                    else store_part(acc, L, grp, r);
                }
                if (kC.G > 1) reduce_parts();
            }
        } else if (vC) {  // stage C, vector unit
            auto cmac = [&](int t, const VLane& L, V& acc) {
                if (isdmrg)
                    vmac<T, OT, true>(W1 + uni(rec->slot[t]) * w1, opc + uni64(rec->opc[t]), gC, L,
                                  (T)rec->alpha[t], acc);
                else
                    vmac<T, OT, false>(W2 + t * w2, opc + uni64(rec->opc[t]), gC, L, (T)1, acc);
            };
            if (pCv) {
                if (wave < tC.total)
                    for (int t = grpV; t < ncs; t += tC.G) cmac(t, LV, accV);
            } else {
                for (int task = wave; task < tC.total; task += NW) {
                    const int ft = task % tC.nft, grp = task / tC.nft;
                    const VLane L = vlane(gC, ft);
                    V acc = V{0, 0, 0, 0};
                    for (int t = grp; t < ncs; t += tC.G) cmac(t, L, acc);
                    if (tC.G == 1) vput<T>(Y, gC, L, acc, true);
                    else vstore_part(acc, grp, ft);
                }
                if (tC.G > 1) vreduce_parts();
            }
        } else if (rbC) {  // stage C, row-blocked (every task owns its output tiles)
            const int base = kCr.tf * kCr.tng;
            for (int task = wave; task < base; task += NW) {
                const RowBlk L = row_blk<T>(gC, (task % kCr.tf) * kTF, task / kCr.tf);
                V acc[kTF];
#pragma unroll
                for (int q = 0; q < kTF; ++q) acc[q] = V{0, 0, 0, 0};
                for (int t0 = 0; t0 < ncs; t0 += kTU) {
                    const Terms<T, OT> R = c_terms<T, OT>(rec, t0, 1, ncs, isdmrg, W1, w1, W2, w2, opc);
                    if (isdmrg) blk_mac_terms<T, OT, true>(R, gC, L, acc);
                    else blk_mac_terms<T, OT, false>(R, gC, L, acc);
                }
                blk_put<T>(Y, gC, L, acc, true);
            }
        } else {  // stage C: Y += sum_t O_c[t] x_c W2[t]  (DMRG: alpha_t O_c[t] x_c W1[slot t])
            for (int task = wave; task < kC.total; task += NW) {
                const int r = task % baseC, grp = task / baseC;
                const RowLane L = row_lane<T>(gC, r % kC.tf, (r / kC.tf) * kTN);
                V acc[kTN];
#pragma unroll
                for (int q = 0; q < kTN; ++q) acc[q] = V{0, 0, 0, 0};
                for (int t0 = grp; t0 < ncs; t0 += kC.G * kTU) {
                    const Terms<T, OT> R = c_terms<T, OT>(rec, t0, kC.G, ncs, isdmrg, W1, w1, W2, w2, opc);
                    if (isdmrg) row_mac_terms<T, OT, true>(R, gC, L, acc);
                    else row_mac_terms<T, OT, false>(R, gC, L, acc);
                }
                if (kC.G == 1) row_put<T>(Y, gC, L, acc, true);
                else store_part(acc, L, grp, r);
            }
            if (kC.G > 1) reduce_parts();
        }
        if (pf) {
            ChunkRec* nx = recs + ((ci + 1 - cb) & 1);
            if (threadIdx.x < pfnt) {
                nx->opb[threadIdx.x] = pfb;
                nx->alpha[threadIdx.x] = pfal;
                nx->slot[threadIdx.x] = pfsl;
            }
            if (threadIdx.x <= pfnc) {
                nx->cbeg[threadIdx.x] = pfg;
                if (threadIdx.x < pfnc) nx->opc[threadIdx.x] = pfc;
            }
            if (threadIdx.x < pfnr) nx->opa[threadIdx.x] = pfa;
        }
        __syncthreads();
    }
    if (pCv) {  // the pair's vector-unit stage C sums into Y
        if (tC.G == 1) {
            if (wave < tC.total) vput<T>(Y, gC, LV, accV, true);
        } else {
            if (wave < tC.total) vstore_part(accV, grpV, wave % tC.nft);
            vreduce_parts();
            __syncthreads();
        }
    }
}

// One workgroup owns one output block at a time (largest first) and
// accumulates every contributing term into it: no atomics, a fixed
// summation order.  Where the pair's input block and stage outputs live
// depends on what fits in LDS (MODE):
//   kModeLds     input block copied into LDS, stage outputs in LDS;
//   kModeLdsW    input block read from global memory, stage outputs in LDS;
//   kModeGlobal  stage outputs in the workgroup's global scratch region.
enum { kModeLds = 0, kModeLdsW = 1, kModeGlobal = 2 };
// Units of several objects can share one launch (MULTI, batched matvec):
// every unit names the object it belongs to (DevBlock::ref), whose lists,
// operator arenas and buffers come from refs[ref].  A single object's launch
// takes its view by value (`one`, in kernel-argument registers for the whole
// launch).  An object's input and output vectors come from its view, or
// from x and y where the view leaves them null.
// Two waves per SIMD in either grid shape: at most 256 registers per lane.
template <typename T, typename OT, int POOL, int NT, int MODE, bool MULTI>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))
    heff_matvec_kernel(const DevBlock* __restrict__ blocks, int nblocks, const HeffRef one,
                       const HeffRef* __restrict__ refs, const T* __restrict__ x,
                       T* __restrict__ y, int* __restrict__ next, const int* __restrict__ seg,
                       int nseg) {
    constexpr int kStage = MODE == kModeGlobal ? 0 : POOL;
    __shared__ T lds[kStage + NT / 64 * 256];
    __shared__ ChunkRec rec[2];
    __shared__ int cur_block;
    T* part = lds + kStage;
    // Units are taken from shared counters, one per segment of the unit
    // list: a workgroup drains the segment of its own XCD first
    // (workgroups are dispatched to the XCDs round-robin, so blockIdx mod 8
    // is the XCD when there are 8 segments), then the others in turn; it
    // leaves once every segment is drained.  One segment: one shared
    // longest-first list.
    const int home = (int)(blockIdx.x % (unsigned)nseg);
    for (;;) {
        if (threadIdx.x == 0) {
            int u = nblocks;
            for (int k = 0; k < nseg; ++k) {
                const int sg = home + k < nseg ? home + k : home + k - nseg;
                const int i = atomicAdd(next + sg, 1);
                if (i < seg[sg + 1] - seg[sg]) {
                    u = seg[sg] + i;
                    break;
                }
            }
            cur_block = u;
        }
        __syncthreads();
        const int bi = uni(cur_block);  // uniform: the unit and its object load through scalar loads
        __syncthreads();
        if (bi >= nblocks) break;
        const DevBlock B = blocks[bi];
        const HeffRef& R = MULTI ? refs[uni(B.ref)] : one;
        const DevPair* __restrict__ pairs = R.pairs;
        const DevChunk* __restrict__ chunks = R.chunks;
        const int* __restrict__ runs = R.runs;
        const DevTerm* __restrict__ terms = R.terms;
        const OT* __restrict__ op0 = static_cast<const OT*>(R.ops[0]);
        const OT* __restrict__ op1 = static_cast<const OT*>(R.ops[1]);
        const OT* __restrict__ op2 = static_cast<const OT*>(R.ops[2]);
        const T* __restrict__ xb = R.x ? static_cast<const T*>(R.x) : x;
        T* __restrict__ Y =
            (B.to_part ? static_cast<T*>(R.partial) : (R.y ? static_cast<T*>(R.y) : y)) + B.y_off;
        const int isdmrg = uni(R.isdmrg), m0 = uni(R.map[0]), m1 = uni(R.map[1]),
                  m2 = uni(R.map[2]);
        constexpr bool FZ = kFused<T> && MODE != kModeGlobal;
        for (int pi = B.pair_begin; pi < B.pair_end; ++pi) {
            const DevPair P = pairs[pi];
            if (MODE == kModeLdsW) {
                pair_chunks<T, OT, NT, FZ>(P, B.d, chunks, runs, terms, op0, op1, op2, xb + P.x_off,
                                           lds, Y, part, rec, isdmrg, m0, m1, m2);
            } else if (MODE == kModeGlobal) {
                T* gw1 = static_cast<T*>(R.scratch) + (int64_t)blockIdx.x * R.scratch_stride;
                pair_chunks<T, OT, NT, FZ>(P, B.d, chunks, runs, terms, op0, op1, op2, xb + P.x_off,
                                           gw1, Y, part, rec, isdmrg, m0, m1, m2);
            } else {
                const T* X = xb + P.x_off;
                const int xsz = P.d[0] * P.d[1] * P.d[2];
                for (int i = threadIdx.x; i < xsz; i += NT) lds[i] = X[i];
                __syncthreads();
                pair_chunks<T, OT, NT, FZ>(P, B.d, chunks, runs, terms, op0, op1, op2, lds,
                                           lds + P.w1_off, Y, part, rec, isdmrg, m0, m1, m2);
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kDiagThreads)
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

// dst = BF16(src * scale) with round to nearest even (the upper 16 bits of
// the FP32 value).
__global__ void to_bf16_kernel(const double* __restrict__ src, const float* __restrict__ scale,
                               uint16_t* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t u = __float_as_uint((float)(src[i] * (double)scale[i]));
        u += 0x7fffu + ((u >> 16) & 1u);
        dst[i] = (uint16_t)(u >> 16);
    }
}

// Output blocks split over several work units: sum their partial copies
// in a fixed order.
template <typename T>
__global__ void split_reduce_kernel(const DevSplit* __restrict__ splits, int nsplits,
                                    const T* __restrict__ partial, T* __restrict__ y) {
    for (int si = blockIdx.x; si < nsplits; si += gridDim.x) {
        const DevSplit S = splits[si];
        for (int i = threadIdx.x; i < S.size; i += blockDim.x) {
            T acc = 0;
            for (int p = 0; p < S.nparts; ++p) acc += partial[S.part_off + (int64_t)p * S.size + i];
            y[S.y_off + i] = acc;
        }
    }
}

inline int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}
}  // namespace

namespace {
// Device cost of contracting one pair in order `ord` (operator legs in
// contraction order), in 16x16x4 matrix-core steps: every stage product
// padded to 16 x 16 output tiles and k-steps of 4 (stages whose output leg
// has at most 4 states run on the vector unit: 64 fibers per wave, one FMA
// per output state and k, a quarter of a step's work), times the products
// of the stage (first-stage operators, terms, stage-C groups), plus a fixed
// price per chunk of terms (the barriers and memory latency of a chunk,
// measured at about a thousand steps' worth) with chunks as the LDS pool
// allows them.
double order_cost(const int (&din)[3], const int (&dout)[3], const int (&ord)[3], int nst,
                  const double (&nprod)[3], double nterms, int pool, double chunk_steps) {
    double cur[3] = {(double)din[0], (double)din[1], (double)din[2]};
    double steps = 0, wsz[2] = {0, 0};
    for (int s = 0; s < nst; ++s) {
        const int leg = ord[s];
        const double K = cur[leg], dn = dout[leg];
        double F = 1;
        for (int k = 0; k < 3; ++k)
            if (k != leg) F *= cur[k];
        const double c = dn <= 4 ? std::ceil(F / 64) * K * dn / 4.0
                                 : std::ceil(F / 16) * std::ceil(dn / 16) * std::ceil(K / 4);
        steps += c * nprod[s];
        cur[leg] = dn;
        if (s < 2) wsz[s] = cur[0] * cur[1] * cur[2];
    }
    const double xsz = (double)din[0] * din[1] * din[2];
    const double avail = std::max(1.0, pool - xsz);
    const double need = nprod[0] * wsz[0] + (nst == 3 ? nterms * wsz[1] : 0.0);
    const double chunks = std::max(std::ceil(need / avail), std::ceil(nterms / kMaxChunkTerms));
    return steps + chunk_steps * chunks;
}

// The same for a three-leg order whose middle stage is fused into stage C
// (middle leg of at most kFuseN states in and out): stage A over the
// distinct first operators, stage C over the distinct last operators (terms
// sharing one are summed in its A operands), the A operands built on the
// fly from every term (`fuse_over`: that work per fiber tile and k-step, in
// matrix-core steps), no middle stage; a chunk holds its first-stage
// outputs and kFuseStage staged values per term.  Returns HUGE_VAL where
// the form does not apply.
double order_cost_fused(const int (&din)[3], const int (&dout)[3], const int (&ord)[3],
                        double nfirst, double nlast, double nterms, int pool, double chunk_steps,
                        double fuse_over) {
    const int a = ord[0], m = ord[1], c = ord[2];
    if (din[m] > kFuseN || dout[m] > kFuseN || dout[c] <= kVN) return HUGE_VAL;
    double cur[3] = {(double)din[0], (double)din[1], (double)din[2]};
    auto fibers = [&](int leg) {
        double F = 1;
        for (int k = 0; k < 3; ++k)
            if (k != leg) F *= cur[k];
        return F;
    };
    double F = fibers(a), K = cur[a], dn = dout[a];
    const double cA = dn <= 4 ? std::ceil(F / 64) * K * dn / 4.0
                              : std::ceil(F / 16) * std::ceil(dn / 16) * std::ceil(K / 4);
    cur[a] = dn;
    const double w1 = cur[0] * cur[1] * cur[2];
    cur[m] = dout[m];
    F = fibers(c);
    K = cur[c];
    dn = dout[c];
    const double cC = std::ceil(F / 16) * std::ceil(dn / 16) * std::ceil(K / 4);
    const double cOp = std::ceil(F / 16) * std::ceil(K / 4) * fuse_over;
    const double xsz = (double)din[0] * din[1] * din[2];
    if (w1 + kFuseStage > pool) return HUGE_VAL;
    const double avail = std::max(1.0, pool - (xsz + w1 + kFuseStage <= pool ? xsz : 0.0));
    const double need = nfirst * w1 + nterms * kFuseStage;
    const double chunks = std::max(std::ceil(need / avail), std::ceil(nterms / kMaxChunkTerms));
    return cA * nfirst + cC * nlast + cOp * nterms + chunk_steps * chunks;
}

// Picks P.ord among the contraction orders (6 for three legs, 2 for DMRG)
// by order_cost (and, with `fuse`, the fused-middle form by
// order_cost_fused); terms: the pair's terms (distinct operator blocks per
// leg give the first-stage and stage-C product counts).
void choose_order(DevPair& P, const int (&dout)[3], const DevTerm* t, int nt, int isdmrg, int pool,
                  bool grouped, double chunk_steps, bool fuse, double fuse_over) {
    double distinct[3] = {0, 0, 0};
    std::vector<int64_t> v(nt);
    for (int k = 0; k < (isdmrg ? 2 : 3); ++k) {
        for (int i = 0; i < nt; ++i) v[i] = t[i].op[k];
        std::sort(v.begin(), v.end());
        distinct[k] = (double)(std::unique(v.begin(), v.end()) - v.begin());
    }
    double best = HUGE_VAL;
    int bo[3] = {P.ord[0], P.ord[1], P.ord[2]};
    int bf = 0;
    const int norders = isdmrg ? 2 : 6;
    for (int o = 0; o < norders; ++o) {
        int ord[3];
        if (isdmrg) {
            ord[0] = o;
            ord[1] = 1 - o;
            ord[2] = 2;
        } else {
            for (int k = 0; k < 3; ++k) ord[k] = hOrd[o][k];
        }
        const int last = isdmrg ? ord[1] : ord[2];
        const double nprod[3] = {distinct[ord[0]], isdmrg ? (double)nt : (double)nt,
                                 grouped && !isdmrg ? distinct[last] : (double)nt};
        const double c = order_cost(P.d, dout, ord, isdmrg ? 2 : 3, nprod, nt, pool, chunk_steps);
        if (c < best * (1 - 1e-12)) {
            best = c;
            bf = 0;
            for (int k = 0; k < 3; ++k) bo[k] = ord[k];
        }
        if (fuse && !isdmrg) {
            const double cf = order_cost_fused(P.d, dout, ord, distinct[ord[0]], distinct[ord[2]],
                                               nt, pool, chunk_steps, fuse_over);
            if (cf < best * (1 - 1e-12)) {
                best = cf;
                bf = 1;
                for (int k = 0; k < 3; ++k) bo[k] = ord[k];
            }
        }
    }
    for (int k = 0; k < 3; ++k) P.ord[k] = bo[k];
    P.fuse = bf;
}
}  // namespace

HeffTuning HeffTuning::from_env() {
    HeffTuning t;
    if (const char* cs = std::getenv("T3NS_HEFF_CHUNK_STEPS")) t.chunk_steps = std::atof(cs);
    if (const char* fv = std::getenv("T3NS_HEFF_FUSE")) t.fuse = std::atoi(fv) != 0;
    if (const char* fo = std::getenv("T3NS_HEFF_FUSE_OVER")) t.fuse_over = std::atof(fo);
    if (const char* wv = std::getenv("T3NS_HEFF_WPC")) t.wpc = std::atoi(wv);
    if (const char* sv = std::getenv("T3NS_HEFF_SCHED")) {
        const char* names[4] = {"cost", "block", "xcd", "xcdblock"};
        t.sched = -1;
        for (int k = 0; k < 4; ++k)
            if (!std::strcmp(sv, names[k])) t.sched = k;
        if (t.sched < 0) throw std::runtime_error("T3NS_HEFF_SCHED must be cost, block, xcd or xcdblock");
    }
    return t;
}

HeffDevice::HeffDevice(const HeffInput& in, int precision, hipStream_t stream,
                       const HeffTuning* tuning)
    : prec_(precision) {
    plan_ = build_heff_plan(in);
    isdmrg_ = plan_.isdmrg;
    for (int k = 0; k < 3; ++k) map_[k] = plan_.map[k];
    size_ = in.site->size();
    const int nops = isdmrg_ ? 2 : 3;
    const InstructionSet& is = *in.iset;
    tune_ = tuning ? *tuning : HeffTuning::from_env();
    const bool f32 = prec_ != PREC_FP64;
    const bool bf16 = prec_ == PREC_BF16;
    // BF16: every operator block is stored divided by a power of two at or
    // above its largest magnitude; the scales of a term's three blocks go
    // into its prefactor (exact, powers of two).
    std::vector<std::vector<std::vector<float>>>& bscale = bscale_;
    bscale.assign(3, {});
    if (bf16) {
        for (int k = 0; k < nops; ++k) {
            const ROperators& op = *in.ops[k];
            bscale[k].resize(op.nrops);
            for (int o = 0; o < op.nrops; ++o) {
                const int nb = (int)op.beginblock[o].size() - 1;
                bscale[k][o].assign(std::max(nb, 0), 1.0f);
                const double* src = (int64_t)op.tel.size() == op.total_size()
                                        ? op.tel.data() + op.op_offset[o]
                                        : (o < (int)op.tel_src.size() ? op.tel_src[o] : nullptr);
                if (!src) continue;
                for (int b = 0; b < nb; ++b) {
                    double m = 0;
                    for (int64_t e = op.beginblock[o][b]; e < op.beginblock[o][b + 1]; ++e)
                        m = std::max(m, std::fabs(src[e]));
                    int ex = 0;
                    if (m > 0) std::frexp(m, &ex);  // m < 2^ex
                    bscale[k][o][b] = (float)std::ldexp(1.0, ex);
                }
            }
        }
    }
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

    dblocks_.upload(hdb, stream);
    dterms_.upload(hdt, stream);
    build_lists(in, stream);
    for (int k = 0; k < nops; ++k) {
        const ROperators& op = *in.ops[k];
        stats_.op_elems[k] = op.total_size();
        if ((int64_t)op.tel.size() == op.total_size()) {
            ops64_[k].upload(op.tel, stream);
        } else if (const ResidentRef ref = resident_find(op)) {  // made on the device
            ops64_[k].alloc(op.total_size());
            T3_HIP_CHECK(hipMemcpyAsync(ops64_[k].get(), ref.ptr(), op.total_size() * sizeof(double),
                                        hipMemcpyDeviceToDevice, stream));
            T3_HIP_CHECK(hipStreamSynchronize(stream));  // before `ref` lets the copy go
        } else if ((int)op.tel_src.size() == op.nrops) {  // straight from the caller's arrays
            ops64_[k].alloc(op.total_size());
            // Small operators are gathered into host chunks (one copy per
            // chunk: a copy call per operator cost more than the data on
            // objects with hundreds of small operators); large ones are
            // copied as they are.  The arena holds the operators in order.
            constexpr int64_t kDirect = 1 << 15, kChunk = 1 << 22;  // elements
            std::vector<double> stage;
            int64_t base = 0;  // arena offset of stage[0]
            auto flush = [&]() {
                if (stage.empty()) return;
                T3_HIP_CHECK(hipMemcpyAsync(ops64_[k].get() + base, stage.data(),
                                            stage.size() * sizeof(double), hipMemcpyHostToDevice,
                                            stream));
                T3_HIP_CHECK(hipStreamSynchronize(stream));  // stage is reused
                stage.clear();
            };
            for (int o = 0; o < op.nrops; ++o) {
                const int64_t off = op.op_offset[o], n = op.op_offset[o + 1] - off;
                if (!n) continue;
                if (n >= kDirect) {
                    T3_HIP_CHECK(hipMemcpyAsync(ops64_[k].get() + off, op.tel_src[o],
                                                n * sizeof(double), hipMemcpyHostToDevice, stream));
                    continue;
                }
                if (!stage.empty() &&
                    (off != base + (int64_t)stage.size() || (int64_t)stage.size() + n > kChunk))
                    flush();
                if (stage.empty()) base = off;
                stage.insert(stage.end(), op.tel_src[o], op.tel_src[o] + n);
            }
            flush();
        } else {
            throw std::runtime_error("HeffDevice: operator host data missing");
        }
    }
    if (prec_ == PREC_FP32) {
        for (int k = 0; k < nops; ++k) {
            const int64_t n = stats_.op_elems[k];
            ops32_[k].alloc(n);
            if (n)
                hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0,
                                   stream, ops64_[k].get(), ops32_[k].get(), n);
        }
    }
    if (bf16) {
        // per-element scale of the BF16 arena: the reciprocal block scale,
        // applied by the conversion kernel (block ranges cover the arena)
        for (int k = 0; k < nops; ++k) {
            const ROperators& op = *in.ops[k];
            const int64_t n = stats_.op_elems[k];
            std::vector<float> inv(std::max<int64_t>(n, 1), 1.0f);
            for (int o = 0; o < op.nrops; ++o)
                for (size_t b = 0; b + 1 < op.beginblock[o].size(); ++b)
                    for (int64_t e = op.beginblock[o][b]; e < op.beginblock[o][b + 1]; ++e)
                        inv[op.op_offset[o] + e] = 1.0f / bscale[k][o][b];
            DevBuf<float> dinv;
            dinv.upload(inv, stream);
            ops16_[k].alloc(std::max<int64_t>(n, 1));
            if (n)
                hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, stream,
                                   ops64_[k].get(), dinv.get(), ops16_[k].get(), n);
            T3_HIP_CHECK(hipStreamSynchronize(stream));
        }
    }
    if (f32) {
        x32_.alloc(size_);
        y32_.alloc(size_);
    }
    T3_HIP_CHECK(hipGetLastError());
    T3_HIP_CHECK(hipStreamSynchronize(stream));
    build_ref();
}

HeffDevice::HeffDevice(const HeffDevice& base, const HeffInput& in, int precision,
                       hipStream_t stream)
    : prec_(precision) {
    if (base.prec_ != PREC_FP64 || precision != PREC_FP32)
        throw std::invalid_argument("a derived Heff is FP32 on an FP64 one");
    plan_ = base.plan_;
    tune_ = base.tune_;
    bscale_.assign(3, {});
    isdmrg_ = base.isdmrg_;
    for (int k = 0; k < 3; ++k) map_[k] = base.map_[k];
    size_ = base.size_;
    ndblocks_ = base.ndblocks_;
    dblocks_.copy_from(base.dblocks_, stream);
    dterms_.copy_from(base.dterms_, stream);
    build_lists(in, stream);
    for (int k = 0; k < 3; ++k) {
        stats_.op_elems[k] = base.stats_.op_elems[k];
        ops64_[k].copy_from(base.ops64_[k], stream);
        const int64_t n = stats_.op_elems[k];
        ops32_[k].alloc(n);
        if (n)
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0,
                               stream, ops64_[k].get(), ops32_[k].get(), n);
    }
    x32_.alloc(size_);
    y32_.alloc(size_);
    T3_HIP_CHECK(hipGetLastError());
    T3_HIP_CHECK(hipStreamSynchronize(stream));
    build_ref();
}

// Work lists of the current tuning (everything the matvec kernels read
// besides the vectors and operator arenas).  Rebuilt by retune().  A
// tuning that leaves the grid shape open (wpc 0) takes the most workgroups
// per CU whose pool cuts the pairs into hardly more chunks than the whole
// pool would (the chunks' LDS demand, re-cut for each pool, grows by at
// most a tenth) with no pair losing its LDS placement.  The lists are
// built first for 4 workgroups per CU on small objects (where that is the
// usual answer) and for 1 on the others, and built again only when the
// estimate picks another shape.  Smaller workgroups wait on memory and
// barriers independently, which pays where chunks are short; where they
// are long the whole pool keeps them fewer.
void HeffDevice::build_lists(const HeffInput& in, hipStream_t stream) {
    constexpr int64_t kSmallObject = 20000;  // vector elements
    const int w0 = tune_.wpc ? tune_.wpc : (size_ < kSmallObject ? 4 : 1);
    const int w = build_lists_at(in, stream, w0, tune_.wpc == 0);
    if (w != w0) build_lists_at(in, stream, w, false);
}

// Builds and uploads the lists for `wpc` workgroups per CU and returns
// wpc; with `pick`, returns the shape to build instead (uploading nothing)
// when the estimate above prefers another.
int HeffDevice::build_lists_at(const HeffInput& in, hipStream_t stream, int wpc, bool pick) {
    drop_graphs();  // captured launches point at the old lists
    {
        const HeffStats old = stats_;
        stats_ = HeffStats{};
        for (int k = 0; k < 3; ++k) stats_.op_elems[k] = old.op_elems[k];
    }
    const int nops = isdmrg_ ? 2 : 3;
    const InstructionSet& is = *in.iset;
    // FP32 and BF16 share the FP32 arithmetic (and its pool and term
    // grouping).
    const bool f32 = prec_ != PREC_FP64;
    const bool bf16 = prec_ == PREC_BF16;
    if (wpc != 1 && wpc != 2 && wpc != 4)
        throw std::invalid_argument("workgroups per CU must be 0 (by structure), 1, 2 or 4");
    wpc_ = wpc;
    int pool = f32 ? (wpc_ == 4 ? Grid<float, 4>::POOL : wpc_ == 2 ? Grid<float, 2>::POOL : kPool32)
                   : (wpc_ == 4 ? Grid<double, 4>::POOL
                                : wpc_ == 2 ? Grid<double, 2>::POOL : kPool64);
    const int chunk_terms = max_chunk_terms(kThreads / wpc_);
    // Input blocks are read from global memory (through L2) by default,
    // which leaves the whole pool to the stage outputs: fewer, larger chunks
    // and no staging barrier, faster on every workload measured (QC-24
    // D = 512 step 370 -> 356 ms at two workgroups per CU; the QC-24
    // D = 128 sweep's matvecs 285 -> 263 ms).  T3NS_HEFF_PLACEMENT=lds
    // copies each input block into the pool when it fits beside one run
    // and one term; =global keeps every stage output in global scratch as
    // well (the path of pairs too large for the pool, forced on any
    // problem for tests); =ldsw is the default.
    bool x_in_lds = false;
    if (const char* pl = std::getenv("T3NS_HEFF_PLACEMENT")) {
        if (!std::strcmp(pl, "lds")) {
            x_in_lds = true;
        } else if (!std::strcmp(pl, "global")) {
            pool = 0;
        } else if (std::strcmp(pl, "ldsw") != 0) {
            throw std::runtime_error("T3NS_HEFF_PLACEMENT must be lds, ldsw or global");
        }
    }
    // Contraction order of each pair: the reference's (fewest flops,
    // src/Heff.c:629-666) or, by default, the one a device cost model
    // prefers (T3NS_HEFF_ORDER=reference|model).
    bool order_by_model = true;
    if (const char* ov = std::getenv("T3NS_HEFF_ORDER")) {
        if (!std::strcmp(ov, "reference")) order_by_model = false;
        else if (std::strcmp(ov, "model") != 0)
            throw std::runtime_error("T3NS_HEFF_ORDER must be reference or model");
    }
    // cost-model prices of this build (HeffTuning)
    const double chunk_steps = tune_.chunk_steps;
    const bool fuse = order_by_model && tune_.fuse;
    const double fuse_over = tune_.fuse_over;
    // Global chunk capacity: at least one stage output of the largest pair.
    int64_t gcap = kGlobalChunkElems;
    for (int sb = 0; sb < plan_.nrblocks; ++sb)
        for (int pi = plan_.pair_begin[sb]; pi < plan_.pair_begin[sb + 1]; ++pi) {
            const PlanPair& pp = plan_.pairs[pi];
            int w[3] = {plan_.dims[pp.oldsb][0], plan_.dims[pp.oldsb][1], plan_.dims[pp.oldsb][2]};
            for (int s = 0; s < (isdmrg_ ? 1 : 2); ++s) {
                const int o = isdmrg_ ? (pp.bestorder == 0 ? 0 : 1) : hOrd[pp.bestorder][s];
                w[o] = plan_.dims[sb][o];
                gcap = std::max<int64_t>(gcap, (int64_t)w[0] * w[1] * w[2]);
            }
        }
    // Work lists per output block, built on host threads: every output
    // block's pairs, chunks, run leads and terms go to its own lists with
    // indices local to them, then the lists are joined in block order with
    // their indices rebased, so the result does not depend on the threads.
    struct Lists {
        std::vector<DevPair> hp;
        std::vector<DevChunk> hc;
        std::vector<int> hr;
        std::vector<DevTerm> ht;
        std::vector<double> pcost;
        double cost = 0;
        HeffStats st;
        DevBlock B{};
        // chunks re-cut for the pools of 1, 2 and 4 workgroups per CU (pick)
        double est[3] = {0, 0, 0};
        int64_t lost[3] = {0, 0, 0};  // pairs that lose the placement the whole pool gives
    };
    const int qpool[3] = {f32 ? kPool32 : kPool64, f32 ? Grid<float, 2>::POOL : Grid<double, 2>::POOL,
                          f32 ? Grid<float, 4>::POOL : Grid<double, 4>::POOL};
    const int qterms[3] = {max_chunk_terms(Grid<double, 1>::NT), max_chunk_terms(Grid<double, 2>::NT),
                           max_chunk_terms(Grid<double, 4>::NT)};
    std::vector<Lists> lists(plan_.nrblocks);
    auto build_block = [&](int sb) {
        Lists& L = lists[sb];
        DevBlock B{};
        B.y_off = plan_.beginblock[sb];
        for (int k = 0; k < 3; ++k) B.d[k] = plan_.dims[sb][k];
        B.pair_begin = (int)L.hp.size();
        for (int pi = plan_.pair_begin[sb]; pi < plan_.pair_begin[sb + 1]; ++pi) {
            const PlanPair& pp = plan_.pairs[pi];
            DevPair P{};
            P.x_off = plan_.beginblock[pp.oldsb];
            for (int k = 0; k < 3; ++k) P.d[k] = plan_.dims[pp.oldsb][k];
            if (isdmrg_) {
                P.ord[0] = pp.bestorder == 0 ? 0 : 1;
                P.ord[1] = 1 - P.ord[0];
                P.ord[2] = 2;
            } else {
                for (int k = 0; k < 3; ++k) P.ord[k] = hOrd[pp.bestorder][k];
            }
            const size_t tb = L.ht.size();
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
                    if (bf16)
                        for (int k = 0; k < nops; ++k)
                            t.alpha *= bscale_[k][ins.instr[k]][e.sbops[k]];
                    L.ht.push_back(t);
                }
            }
            if ((int64_t)L.ht.size() > INT32_MAX)
                throw std::runtime_error("HeffDevice: more than 2^31 terms");
            if (L.ht.size() == tb) continue;
            if (order_by_model)
                choose_order(P, B.d, L.ht.data() + tb, (int)(L.ht.size() - tb), isdmrg_, pool, f32,
                             chunk_steps, fuse, fuse_over);
            // Terms sorted by their first-stage operator block: runs of equal
            // first operators share one first-stage output.
            const int first = P.ord[0];
            std::stable_sort(L.ht.begin() + tb, L.ht.end(), [first](const DevTerm& a, const DevTerm& b) {
                return a.op[first] < b.op[first];
            });
            // stage output sizes
            int w[3] = {P.d[0], P.d[1], P.d[2]};
            for (int s = 0; s < (isdmrg_ ? 1 : 2); ++s) {
                const int o = P.ord[s];
                w[o] = B.d[o];
                const int64_t sz = (int64_t)w[0] * w[1] * w[2];
                if (s == 0) P.w1 = (int)sz;
                else P.w2 = (int)sz;
            }
            const int64_t xsz = (int64_t)P.d[0] * P.d[1] * P.d[2];
            // LDS when one run and one term fit next to the input block
            // (kModeLds) or without it (kModeLdsW)
            const int64_t wmin = (int64_t)P.w1 + (isdmrg_ ? 0 : (P.fuse ? kFuseStage : P.w2));
            P.xlds = x_in_lds && xsz + wmin <= pool;
            P.wlds = wmin <= pool;
            P.w1_off = P.xlds ? (int)xsz : 0;
            const int64_t avail = pool - P.w1_off;
            // chunks
            P.chunk_begin = (int)L.hc.size();
            DevChunk C{};
            int nruns = 0, nterms = 0;
            int64_t prev = -1;
            // last operators of the chunk's groups.  Terms are grouped on the
            // FP32 path and in fused FP64 pairs; other FP64 pairs and DMRG
            // objects keep one group per term (the FP64 kernel's stage B has
            // no register room for the group loops).
            const bool group = !isdmrg_ && (f32 || P.fuse);
            const int last = isdmrg_ ? P.ord[1] : P.ord[2];
            std::vector<int64_t> cops;
            auto group_of = [&](const DevTerm& tm) {
                if (!group) return -1;
                for (size_t g = 0; g < cops.size(); ++g)
                    if (cops[g] == tm.op[last]) return (int)g;
                return -1;
            };
            auto close = [&](int end) {
                C.term_end = end;
                C.run_end = (int)L.hr.size();
                const int nt = end - C.term_begin;
                C.ncs = group ? (int)cops.size() : nt;
                // terms in group order (stable: each group keeps the order
                // of its terms); run leads re-pointed at a term of the run
                std::stable_sort(L.ht.begin() + C.term_begin, L.ht.begin() + end,
                                 [](const DevTerm& x, const DevTerm& y) { return x.cslot < y.cslot; });
                for (int r = 0; r < nruns; ++r) L.hr[C.run_begin + r] = -1;
                for (int t = C.term_begin; t < end; ++t)
                    if (L.hr[C.run_begin + L.ht[t].slot] < 0) L.hr[C.run_begin + L.ht[t].slot] = t;
                for (int g = 0, t = 0; g <= C.ncs; ++g) {
                    while (t < nt && L.ht[C.term_begin + t].cslot < g) ++t;
                    L.hr.push_back(t);
                }
                const int64_t r1 = (int64_t)nruns * P.w1;
                C.w2_off = P.wlds ? (int)r1 : (int)gcap;
                L.hc.push_back(C);
            };
            for (size_t t = tb; t < L.ht.size(); ++t) {
                const bool newrun = nterms == 0 || L.ht[t].op[first] != prev;
                const bool newc = !group || nterms == 0 || group_of(L.ht[t]) < 0;
                const int64_t r1 = (int64_t)(nruns + newrun) * P.w1;
                const int64_t r2 =
                    isdmrg_ ? 0
                    : P.fuse ? (int64_t)(nterms + 1) * kFuseStage
                             : (int64_t)((nterms == 0 ? 0 : (group ? (int)cops.size() : nterms)) + newc) *
                                   P.w2;
                const bool fits = nterms < chunk_terms &&
                                  (P.wlds ? r1 + r2 <= avail : r1 <= gcap && r2 <= gcap);
                if (nterms > 0 && !fits) {
                    close((int)t);
                    nruns = nterms = 0;
                }
                if (nterms == 0) {
                    C = DevChunk{};
                    C.term_begin = (int)t;
                    C.run_begin = (int)L.hr.size();
                    cops.clear();
                }
                if (nterms == 0 || L.ht[t].op[first] != prev) {
                    L.hr.push_back((int)t);
                    ++nruns;
                }
                L.ht[t].slot = nruns - 1;
                if (!group) {
                    L.ht[t].cslot = nterms;
                } else {
                    int g = group_of(L.ht[t]);
                    if (g < 0) {
                        g = (int)cops.size();
                        cops.push_back(L.ht[t].op[last]);
                    }
                    L.ht[t].cslot = g;
                }
                prev = L.ht[t].op[first];
                ++nterms;
            }
            close((int)L.ht.size());
            P.chunk_end = (int)L.hc.size();
            {
                // matrix-core steps: tiles x ceil(K/4) per stage output
                int cd[3] = {P.d[0], P.d[1], P.d[2]};
                double st[3];
                for (int s2 = 0; s2 < (isdmrg_ ? 2 : 3); ++s2) {
                    const int o = P.ord[s2];
                    double F = 1;
                    for (int k = 0; k < 3; ++k)
                        if (k != o) F *= cd[k];
                    st[s2] = std::ceil(F / 16) * std::ceil(B.d[o] / 16.0) * std::ceil(cd[o] / 4.0);
                    cd[o] = B.d[o];
                }
                double nr = 0, ntt = 0, ncs = 0;
                for (int ci = P.chunk_begin; ci < P.chunk_end; ++ci) {
                    nr += L.hc[ci].run_end - L.hc[ci].run_begin;
                    ntt += L.hc[ci].term_end - L.hc[ci].term_begin;
                    ncs += L.hc[ci].ncs;
                }
                L.st.stage_steps[0] += nr * st[0];
                L.st.stage_steps[1] += P.fuse ? 0.0 : ntt * st[1];
                L.st.pairs_fused += P.fuse;
                if (!isdmrg_) L.st.stage_steps[2] += ncs * st[2];
                L.st.nruns += (int64_t)nr;
                L.st.pairs_global_x += !P.xlds;
                L.st.pairs_global_w += !P.wlds;
            }
            {
                double c = 0;
                for (int ci = P.chunk_begin; ci < P.chunk_end; ++ci) {
                    const double nr = L.hc[ci].run_end - L.hc[ci].run_begin;
                    const double nt = L.hc[ci].term_end - L.hc[ci].term_begin;
                    c += nr * xsz * B.d[P.ord[0]] + nt * (double)P.w1 * B.d[P.ord[1]] +
                         (isdmrg_ ? 0.0 : L.hc[ci].ncs * (double)P.w2 * B.d[P.ord[2]]);
                }
                L.cost += c;
                L.pcost.push_back(c);
            }
            if (pick) {
                // LDS demand of the pair's chunks (first- and second-stage
                // outputs), spread over the smaller pools
                double dem = 0, ntt = 0;
                for (int ci = P.chunk_begin; ci < P.chunk_end; ++ci) {
                    const DevChunk& Q = L.hc[ci];
                    const double nt = Q.term_end - Q.term_begin;
                    dem += (double)(Q.run_end - Q.run_begin) * P.w1 +
                           (isdmrg_ ? 0.0 : P.fuse ? nt * kFuseStage : (double)Q.ncs * P.w2);
                    ntt += nt;
                }
                const bool w1fit = wmin <= qpool[0], x1fit = x_in_lds && xsz + wmin <= qpool[0];
                for (int q = 0; q < 3; ++q) {
                    const bool xq = x_in_lds && xsz + wmin <= qpool[q];
                    if ((w1fit && wmin > qpool[q]) || (x1fit && !xq)) ++L.lost[q];
                    const double avq = std::max(1.0, (double)qpool[q] - (xq ? (double)xsz : 0.0));
                    L.est[q] += std::max(std::ceil(dem / avq), std::ceil(ntt / qterms[q]));
                }
            }
            L.hp.push_back(P);
        }
        B.pair_end = (int)L.hp.size();
        L.B = B;
    };
    run_blocks_threaded(plan_.nrblocks, build_block, (double)plan_.pairs.size(), 2e4);
    if (pick && pool > 0) {
        // the whole pool's chunk count: counted when these lists are its,
        // else estimated like the others
        double est[3] = {0, 0, 0};
        int64_t lost[3] = {0, 0, 0};
        double nch = 0;
        for (const Lists& L : lists) {
            nch += (double)L.hc.size();
            for (int q = 0; q < 3; ++q) {
                est[q] += L.est[q];
                lost[q] += L.lost[q];
            }
        }
        const double base = wpc == 1 ? nch : est[0];
        int w = 1;
        for (int q = 2; q >= 1 && w == 1; --q)
            if (base > 0 && !lost[q] && est[q] <= 1.1 * base) w = q == 2 ? 4 : 2;
        if (w != wpc) return w;
    }
    std::vector<DevPair> hp;
    std::vector<DevChunk> hc;
    std::vector<int> hr;
    std::vector<DevTerm> ht;
    std::vector<DevBlock> hb(plan_.nrblocks);
    std::vector<double> cost(plan_.nrblocks, 0.0);
    std::vector<double> pcost;
    {
        std::vector<int64_t> pb(plan_.nrblocks + 1, 0), cb(pb), rb(pb), tb(pb);
        for (int sb = 0; sb < plan_.nrblocks; ++sb) {
            const Lists& L = lists[sb];
            pb[sb + 1] = pb[sb] + (int64_t)L.hp.size();
            cb[sb + 1] = cb[sb] + (int64_t)L.hc.size();
            rb[sb + 1] = rb[sb] + (int64_t)L.hr.size();
            tb[sb + 1] = tb[sb] + (int64_t)L.ht.size();
            for (int k = 0; k < 3; ++k) stats_.stage_steps[k] += L.st.stage_steps[k];
            stats_.nruns += L.st.nruns;
            stats_.pairs_global_x += L.st.pairs_global_x;
            stats_.pairs_global_w += L.st.pairs_global_w;
            stats_.pairs_fused += L.st.pairs_fused;
            cost[sb] = L.cost;
        }
        const int64_t nb = plan_.nrblocks;
        if (tb[nb] > INT32_MAX || rb[nb] > INT32_MAX || cb[nb] > INT32_MAX)
            throw std::runtime_error("HeffDevice: more than 2^31 terms");
        hp.resize(pb[nb]);
        hc.resize(cb[nb]);
        hr.resize(rb[nb]);
        ht.resize(tb[nb]);
        pcost.resize(pb[nb]);
        run_blocks_threaded(plan_.nrblocks, [&](int sb) {
            Lists& L = lists[sb];
            DevBlock B = L.B;
            B.pair_begin += (int)pb[sb];
            B.pair_end += (int)pb[sb];
            hb[sb] = B;
            for (size_t i = 0; i < L.hp.size(); ++i) {
                DevPair P = L.hp[i];
                P.chunk_begin += (int)cb[sb];
                P.chunk_end += (int)cb[sb];
                hp[pb[sb] + i] = P;
                pcost[pb[sb] + i] = L.pcost[i];
            }
            for (size_t i = 0; i < L.hc.size(); ++i) {
                DevChunk C = L.hc[i];
                // run leads are term indices; the group offsets after them
                // are chunk-relative
                for (int r = C.run_begin; r < C.run_end; ++r) L.hr[r] += (int)tb[sb];
                C.term_begin += (int)tb[sb];
                C.term_end += (int)tb[sb];
                C.run_begin += (int)rb[sb];
                C.run_end += (int)rb[sb];
                hc[cb[sb] + i] = C;
            }
            std::copy(L.hr.begin(), L.hr.end(), hr.begin() + rb[sb]);
            std::copy(L.ht.begin(), L.ht.end(), ht.begin() + tb[sb]);
            std::vector<DevPair>().swap(L.hp);  // release as we go
            std::vector<DevTerm>().swap(L.ht);
        }, (double)tb[nb], 1e6);
    }
    const int64_t maxscratch = 2 * gcap;
    // Work units: output blocks whose cost exceeds a share of the machine
    // are split into slices of their pairs that accumulate into partial
    // copies, summed afterwards in a fixed order.
    int dev = 0, ncu = 256;
    T3_HIP_CHECK(hipGetDevice(&dev));
    T3_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    double total = 0;
    for (double c : cost) total += c;
    const double unit_cap = std::max(1.0, total / (4.0 * ncu * wpc_));  // 4 units per workgroup
    std::vector<DevBlock> units;
    std::vector<double> ucost;
    std::vector<int> uglob;
    std::vector<int> usb;  // output block of each unit
    std::vector<DevSplit> hs;
    int64_t part_elems = 0;
    for (const DevBlock& B : hb) {
        if (B.pair_end == B.pair_begin) continue;
        const int64_t ysz = (int64_t)B.d[0] * B.d[1] * B.d[2];
        // Pairs whose buffers fit in LDS and pairs that need global scratch
        // go to different units (they run in different kernels).
        std::vector<std::pair<int, int>> slices;
        std::vector<double> scost;
        std::vector<int> sglob;
        double acc = 0;
        int beg = B.pair_begin;
        // kernel of a pair: kModeLds, kModeLdsW or kModeGlobal
        auto is_glob = [&](int pi) { return !hp[pi].wlds ? 2 : (!hp[pi].xlds ? 1 : 0); };
        for (int pi = B.pair_begin; pi < B.pair_end; ++pi) {
            if (pi > beg && (acc + pcost[pi] > unit_cap || is_glob(pi) != is_glob(beg))) {
                slices.push_back({beg, pi});
                scost.push_back(acc);
                sglob.push_back(is_glob(beg));
                beg = pi;
                acc = 0;
            }
            acc += pcost[pi];
        }
        slices.push_back({beg, B.pair_end});
        scost.push_back(acc);
        sglob.push_back(is_glob(beg));
        const int sbi = (int)(&B - hb.data());
        if (slices.size() == 1) {
            units.push_back(B);
            ucost.push_back(scost[0]);
            uglob.push_back(sglob[0]);
            usb.push_back(sbi);
            continue;
        }
        DevSplit S{};
        S.y_off = B.y_off;
        S.part_off = part_elems;
        S.nparts = (int)slices.size();
        S.size = (int)ysz;
        hs.push_back(S);
        for (size_t k = 0; k < slices.size(); ++k) {
            DevBlock U = B;
            U.y_off = part_elems + (int64_t)k * ysz;
            U.pair_begin = slices[k].first;
            U.pair_end = slices[k].second;
            U.to_part = 1;
            units.push_back(U);
            ucost.push_back(scost[k]);
            uglob.push_back(sglob[k]);
            usb.push_back(sbi);
        }
        part_elems += (int64_t)slices.size() * ysz;
    }
    // Units in kernel order (kModeLds, kModeLdsW, kModeGlobal), each list
    // ordered by the schedule (T3NS_HEFF_SCHED):
    //   cost      costliest first, one shared list (dynamic longest-first);
    //   block     output-block order (neighbouring blocks share sectors,
    //             hence operator blocks), the cheapest tenth of the cost
    //             moved to the end, costliest first, so the schedule drains
    //             evenly;
    //   xcd       output-block order cut into one segment of equal cost
    //             per XCD (8), costliest first within a segment; a
    //             workgroup drains its XCD's segment, then the others;
    //   xcdblock  the same segments, each in the block order above.
    const int sched = tune_.sched;
    auto by_cost = [&](std::vector<int>& v) {
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return ucost[a] > ucost[b]; });
    };
    // block order with the cheapest tenth of the cost at the end
    auto by_block = [&](std::vector<int>& v) {
        std::vector<int> c = v;
        std::stable_sort(c.begin(), c.end(), [&](int a, int b) { return ucost[a] < ucost[b]; });
        double tot = 0, acc = 0;
        for (int i : v) tot += ucost[i];
        std::vector<char> tail(units.size(), 0);
        for (int i : c) {
            if (acc + ucost[i] > 0.1 * tot) break;
            acc += ucost[i];
            tail[i] = 1;
        }
        std::vector<int> head, rest;
        for (int i : v) (tail[i] ? rest : head).push_back(i);  // v is in block order
        by_cost(rest);
        v = head;
        v.insert(v.end(), rest.begin(), rest.end());
    };
    std::vector<DevBlock> hbs;
    std::vector<double> hbcost;
    std::vector<int> hseg;
    for (int m = 0; m < 3; ++m) {
        std::vector<int> v;
        for (size_t i = 0; i < units.size(); ++i)
            if (uglob[i] == m) v.push_back((int)i);  // block order
        const int nseg = (sched >= 2 && m == 0 && v.size() >= 64) ? kSegments : 1;
        std::vector<std::vector<int>> segs(nseg);
        double tot = 0;
        for (int i : v) tot += ucost[i];
        double acc = 0;
        for (int i : v) {
            const int sg = std::min(nseg - 1, (int)(acc / tot * nseg));
            segs[sg].push_back(i);
            acc += ucost[i];
        }
        nseg_[m] = nseg;
        const int base = (int)hbs.size();
        for (int sg = 0; sg <= kSegments; ++sg) hseg.push_back(0);
        int* sb = hseg.data() + (kSegments + 1) * m;
        for (int sg = 0; sg < nseg; ++sg) {
            if (sched == 0 || sched == 2) by_cost(segs[sg]);
            else by_block(segs[sg]);
            sb[sg] = (int)hbs.size() - base;
            for (int i : segs[sg]) {
                hbs.push_back(units[i]);
                hbcost.push_back(ucost[i]);
            }
        }
        sb[nseg] = (int)hbs.size() - base;
        (m == 0 ? nfast_ : m == 1 ? nxg_ : nglob_) = (int)hbs.size() - base;
    }
    stats_.nunits = (int64_t)hbs.size();
    stats_.nsplit_blocks = (int64_t)hs.size();
    nsplits_ = (int)hs.size();
    part_elems_ = part_elems;

    stats_.nblocks = (int64_t)hbs.size();
    stats_.npairs = (int64_t)hp.size();
    stats_.nterms = (int64_t)ht.size();
    stats_.size = size_;
    stats_.flops = plan_.flops;
    stats_.scratch_per_wg = maxscratch;
    stats_.nchunks = (int64_t)hc.size();
    stats_.mfma_steps = stats_.stage_steps[0] + stats_.stage_steps[1] + stats_.stage_steps[2];

    blocks_.upload(hbs, stream);
    hunits_ = hbs;
    hucost_ = hbcost;
    pairs_.upload(hp, stream);
    chunks_.upload(hc, stream);
    runs_.upload(hr, stream);
    terms_.upload(ht, stream);
    // Persistent grids: one workgroup per CU; global scratch only for the
    // workgroups of the global-scratch kernel.
    ncu_ = ncu;
    nwg_glob_ = std::max(1, std::min(nglob_, ncu * wpc_));
    if (f32) scratch32_.alloc(nglob_ ? (size_t)nwg_glob_ * maxscratch : 1);
    else scratch64_.alloc(nglob_ ? (size_t)nwg_glob_ * maxscratch : 1);
    counter_.alloc(3 * kSegments);
    seg_.upload(hseg, stream);
    splits_.upload(hs, stream);
    if (f32) part32_.alloc(std::max<int64_t>(1, part_elems_));
    else part64_.alloc(std::max<int64_t>(1, part_elems_));
    T3_HIP_CHECK(hipGetLastError());
    T3_HIP_CHECK(hipStreamSynchronize(stream));
    return wpc;
}

void HeffDevice::retune(const HeffInput& in, const HeffTuning& t, hipStream_t stream) {
    tune_ = t;
    build_lists(in, stream);
    build_ref();
}


HeffDevice::~HeffDevice() { drop_graphs(); }

template <typename T, typename OT, int WPC>
void HeffDevice::launch(const T* x, T* y, T* part, hipStream_t s) {
    constexpr int NT = Grid<T, WPC>::NT, POOL = Grid<T, WPC>::POOL;
    const int nwg = ncu_ * WPC;
    hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(size_)), dim3(256), 0, s, y, size_, (T)0);
    hipLaunchKernelGGL((fill_kernel<int>), dim3(1), dim3(64), 0, s, counter_.get(),
                       (int64_t)(3 * kSegments), 0);
    if (part_elems_)
        hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(part_elems_)), dim3(256), 0, s, part,
                           part_elems_, (T)0);
    if (nfast_)
        hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeLds, false>),
                           dim3(std::min(nfast_, nwg)), dim3(NT), 0, s, blocks_.get(), nfast_,
                           href_, (const HeffRef*)nullptr, x, y, counter_.get(), seg_.get(),
                           nseg_[0]);
    if (nxg_)
        hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeLdsW, false>),
                           dim3(std::min(nxg_, nwg)), dim3(NT), 0, s, blocks_.get() + nfast_,
                           nxg_, href_, (const HeffRef*)nullptr, x, y,
                           counter_.get() + kSegments, seg_.get() + (kSegments + 1), nseg_[1]);
    if (nglob_)
        hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeGlobal, false>),
                           dim3(nwg_glob_), dim3(NT), 0, s, blocks_.get() + nfast_ + nxg_, nglob_,
                           href_, (const HeffRef*)nullptr, x, y, counter_.get() + 2 * kSegments,
                           seg_.get() + 2 * (kSegments + 1), nseg_[2]);
    if (nsplits_)
        hipLaunchKernelGGL((split_reduce_kernel<T>), dim3(std::min(nsplits_, 4096)), dim3(256), 0,
                           s, splits_.get(), nsplits_, part, y);
}

template <typename T, typename OT>
void HeffDevice::launch_any(const T* x, T* y, T* part, hipStream_t s) {
    if (wpc_ == 4) launch<T, OT, 4>(x, y, part, s);
    else if (wpc_ == 2) launch<T, OT, 2>(x, y, part, s);
    else launch<T, OT, 1>(x, y, part, s);
}

// The object's device view for the kernels (input and output left to the
// launch).  Rebuilt whenever lists or arenas are (re)allocated.
void HeffDevice::build_ref() {
    HeffRef r{};
    r.pairs = pairs_.get();
    r.chunks = chunks_.get();
    r.runs = runs_.get();
    r.terms = terms_.get();
    for (int k = 0; k < 3; ++k)
        r.ops[k] = prec_ == PREC_FP64 ? (const void*)ops64_[k].get()
                   : prec_ == PREC_FP32 ? (const void*)ops32_[k].get()
                                        : (const void*)ops16_[k].get();
    r.x = nullptr;
    r.y = nullptr;
    r.partial = prec_ == PREC_FP64 ? (void*)part64_.get() : (void*)part32_.get();
    r.scratch = prec_ == PREC_FP64 ? (void*)scratch64_.get() : (void*)scratch32_.get();
    r.scratch_stride = stats_.scratch_per_wg;
    r.isdmrg = isdmrg_;
    for (int k = 0; k < 3; ++k) r.map[k] = map_[k];
    href_ = r;
}

void HeffDevice::matvec(const double* dvec, double* dres, hipStream_t s) {
    // Graph replay: the launch sequence of one (input, output) pair of
    // device arrays is captured once and replayed (a captured stream only;
    // the legacy default stream cannot be captured).  Every node is a
    // kernel (the fills included), so a replay repeats exactly the eager
    // work.
    if (use_graphs_ && s != nullptr && s != hipStreamPerThread) {
        hipGraphExec_t exec = nullptr;
        for (const auto& g : graphs_)
            if (g.x == dvec && g.y == dres) exec = g.exec;
        if (!exec) {
            hipGraph_t graph = nullptr;
            T3_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueue(dvec, dres, s);
            T3_HIP_CHECK(hipStreamEndCapture(s, &graph));
            T3_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            T3_HIP_CHECK(hipGraphDestroy(graph));
            if (graphs_.size() >= kMaxGraphs) {  // oldest out
                T3_HIP_CHECK(hipGraphExecDestroy(graphs_.front().exec));
                graphs_.erase(graphs_.begin());
            }
            graphs_.push_back({dvec, dres, exec});
        }
        T3_HIP_CHECK(hipGraphLaunch(exec, s));
        return;
    }
    enqueue(dvec, dres, s);
}

void HeffDevice::set_graphs(bool on) {
    use_graphs_ = on;
    if (!on) drop_graphs();
}

void HeffDevice::drop_graphs() {
    for (const auto& g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

void HeffDevice::enqueue(const double* dvec, double* dres, hipStream_t s) {
    if (prec_ == PREC_FP64) {
        launch_any<double, double>(dvec, dres, part64_.get(), s);
    } else {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(size_)), dim3(256), 0, s,
                           dvec, x32_.get(), size_);
        if (prec_ == PREC_BF16)
            launch_any<float, uint16_t>(x32_.get(), y32_.get(), part32_.get(), s);
        else
            launch_any<float, float>(x32_.get(), y32_.get(), part32_.get(), s);
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(size_)), dim3(256), 0, s,
                           y32_.get(), dres, size_);
    }
    T3_HIP_CHECK(hipGetLastError());
}

template <typename T, typename OT>
void HeffBatch::launch(HeffDevice* const* hs, const T* const* x, T* const* y, int n,
                       hipStream_t s) {
    // Units of the kernels with no global scratch from every object, one
    // longest-first list per kernel and grid shape; each unit names its
    // object.
    std::vector<DevBlock> units[kShapes][2];  // [grid shape][kernel]
    std::vector<double> cost[kShapes][2];
    std::vector<HeffRef> refs(n);
    for (int i = 0; i < n; ++i) {
        HeffDevice& h = *hs[i];
        refs[i] = h.href_;
        refs[i].x = x[i];
        refs[i].y = y[i];
        const int g = h.wpc_ == 4 ? 2 : h.wpc_ - 1;
        const int first[2] = {0, h.nfast_}, count[2] = {h.nfast_, h.nxg_};
        for (int m = 0; m < 2; ++m)
            for (int u = first[m]; u < first[m] + count[m]; ++u) {
                DevBlock B = h.hunits_[u];
                B.ref = i;
                units[g][m].push_back(B);
                cost[g][m].push_back(h.hucost_[u]);
            }
    }
    for (int g = 0; g < kShapes; ++g)
        for (int m = 0; m < 2; ++m) {
            std::vector<int> idx(units[g][m].size());
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(),
                             [&](int a, int b) { return cost[g][m][a] > cost[g][m][b]; });
            std::vector<DevBlock> v;
            for (int k : idx) v.push_back(units[g][m][k]);
            units[g][m].swap(v);
        }
    std::vector<DevBlock> all;
    std::vector<int> seg;  // one segment per kernel and grid shape
    for (int g = 0; g < kShapes; ++g)
        for (int m = 0; m < 2; ++m) {
            seg.push_back(0);
            seg.push_back((int)units[g][m].size());
            all.insert(all.end(), units[g][m].begin(), units[g][m].end());
        }
    units_.upload(all, s);
    refs_.upload(refs, s);
    seg_.upload(seg, s);
    counter_.reserve(2 * kShapes);
    hipLaunchKernelGGL((fill_kernel<int>), dim3(1), dim3(64), 0, s, counter_.get(),
                       (int64_t)(2 * kShapes), 0);
    for (int i = 0; i < n; ++i) {
        HeffDevice& h = *hs[i];
        hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(h.size_)), dim3(256), 0, s, y[i],
                           h.size_, (T)0);
        if (h.part_elems_)
            hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(h.part_elems_)), dim3(256), 0, s,
                               static_cast<T*>(h.href_.partial), h.part_elems_, (T)0);
    }
    ncu_ = hs[0]->ncu_;
    int base = 0;
    launch_grid<T, OT, 1>(units[0], base, 0, s);
    base += (int)(units[0][0].size() + units[0][1].size());
    launch_grid<T, OT, 2>(units[1], base, 1, s);
    base += (int)(units[1][0].size() + units[1][1].size());
    launch_grid<T, OT, 4>(units[2], base, 2, s);
    // units that need global scratch run in their objects' own launches
    // (the scratch is sized for the object's grid); then the split sums
    for (int i = 0; i < n; ++i) {
        HeffDevice& h = *hs[i];
        if (h.nglob_) {
            hipLaunchKernelGGL((fill_kernel<int>), dim3(1), dim3(64), 0, s,
                               h.counter_.get() + 2 * HeffDevice::kSegments, (int64_t)1, 0);
            if (h.wpc_ == 4)
                launch_glob<T, OT, 4>(h, x[i], y[i], s);
            else if (h.wpc_ == 2)
                launch_glob<T, OT, 2>(h, x[i], y[i], s);
            else
                launch_glob<T, OT, 1>(h, x[i], y[i], s);
        }
        if (h.nsplits_)
            hipLaunchKernelGGL((split_reduce_kernel<T>), dim3(std::min(h.nsplits_, 4096)), dim3(256),
                               0, s, h.splits_.get(), h.nsplits_, static_cast<T*>(h.href_.partial),
                               y[i]);
    }
    T3_HIP_CHECK(hipGetLastError());
}

// The batch's kernels of one grid shape: units[m] (kernel m) start at
// `base` in the uploaded unit list, counters and segments at 2 shape.
template <typename T, typename OT, int WPC>
void HeffBatch::launch_grid(const std::vector<DevBlock> (&units)[2], int base, int shape,
                            hipStream_t s) {
    constexpr int NT = Grid<T, WPC>::NT, POOL = Grid<T, WPC>::POOL;
    const int nwg = ncu_ * WPC;
    const int n0 = (int)units[0].size(), n1 = (int)units[1].size();
    const int c = 2 * shape;
    if (n0)
        hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeLds, true>),
                           dim3(std::min(n0, nwg)), dim3(NT), 0, s, units_.get() + base, n0,
                           HeffRef{}, refs_.get(), (const T*)nullptr, (T*)nullptr,
                           counter_.get() + c, seg_.get() + 2 * c, 1);
    if (n1)
        hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeLdsW, true>),
                           dim3(std::min(n1, nwg)), dim3(NT), 0, s, units_.get() + base + n0, n1,
                           HeffRef{}, refs_.get(), (const T*)nullptr, (T*)nullptr,
                           counter_.get() + c + 1, seg_.get() + 2 * c + 2, 1);
}

// An object's global-scratch kernel inside a batch (its own grid).
template <typename T, typename OT, int WPC>
void HeffBatch::launch_glob(HeffDevice& h, const T* x, T* y, hipStream_t s) {
    constexpr int NT = Grid<T, WPC>::NT, POOL = Grid<T, WPC>::POOL;
    hipLaunchKernelGGL((heff_matvec_kernel<T, OT, POOL, NT, kModeGlobal, false>),
                       dim3(h.nwg_glob_), dim3(NT), 0, s, h.blocks_.get() + h.nfast_ + h.nxg_,
                       h.nglob_, h.href_, (const HeffRef*)nullptr, x, y,
                       h.counter_.get() + 2 * HeffDevice::kSegments,
                       h.seg_.get() + 2 * (HeffDevice::kSegments + 1), h.nseg_[2]);
}

void HeffBatch::matvec(HeffDevice* const* hs, const double* const* x, double* const* y, int n,
                       hipStream_t s) {
    if (n <= 0) return;
    const int prec = hs[0]->prec_;
    for (int i = 1; i < n; ++i)
        if (hs[i]->prec_ != prec) throw std::invalid_argument("batched matvec: mixed precisions");
    if (prec == PREC_FP64) {
        launch<double, double>(hs, x, y, n, s);
        return;
    }
    std::vector<const float*> x32(n);
    std::vector<float*> y32(n);
    for (int i = 0; i < n; ++i) {
        HeffDevice& h = *hs[i];
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(h.size_)), dim3(256), 0,
                           s, x[i], h.x32_.get(), h.size_);
        x32[i] = h.x32_.get();
        y32[i] = h.y32_.get();
    }
    if (prec == PREC_BF16)
        launch<float, uint16_t>(hs, x32.data(), y32.data(), n, s);
    else
        launch<float, float>(hs, x32.data(), y32.data(), n, s);
    for (int i = 0; i < n; ++i)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(hs[i]->size_)), dim3(256),
                           0, s, y32[i], y[i], hs[i]->size_);
    T3_HIP_CHECK(hipGetLastError());
}

void HeffDevice::diagonal(double* ddiag, hipStream_t s) {
    hipLaunchKernelGGL((fill_kernel<double>), dim3(grid_for(size_)), dim3(256), 0, s, ddiag, size_,
                       0.0);
    if (ndblocks_)
        hipLaunchKernelGGL((heff_diag_kernel<double>), dim3(std::min(ndblocks_, 8192)),
                           dim3(kDiagThreads), 0,
                           s, dblocks_.get(), ndblocks_, dterms_.get(), ops64_[0].get(),
                           ops64_[1].get(), ops64_[2].get(), ddiag, isdmrg_, map_[0], map_[1],
                           map_[2]);
    T3_HIP_CHECK(hipGetLastError());
}

}  // namespace t3
