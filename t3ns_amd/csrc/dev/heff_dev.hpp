// Device-resident effective Hamiltonian: the plan flattened into work lists
// in HBM and the kernels that execute Psi' = Heff Psi and diag(Heff).
// Replaces the OpenMP replay loop of the reference (src/Heff.c:916-1000) and
// make_diagonal (src/Heff.c:1268-1312).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../host/heff_plan.hpp"
#include "hip_util.hpp"

namespace t3 {

enum Precision : int { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16 = 2 };

struct DevBlock {      // one work unit: an output block or a slice of its pairs
    int64_t y_off;     // element offset of the destination (vector or partials)
    int d[3];          // leg dims in operator order
    int pair_begin;
    int pair_end;
    int to_part;       // 1: accumulate into the partial buffer
    int ref;           // the object it belongs to (index into the launch's refs)
};
struct DevPair;
struct DevChunk;
struct DevTerm;
// Device addresses of one object's work lists, operator arenas and buffers
// (the kernels' view of a HeffDevice; a batched launch holds several).
struct HeffRef {
    const DevPair* pairs;
    const DevChunk* chunks;
    const int* runs;
    const DevTerm* terms;
    const void* ops[3];  // arenas in the storage type
    const void* x;       // input vector in the arithmetic type (null: the launch's)
    void* y;             // output vector (null: the launch's)
    void* partial;       // partial copies of split output blocks
    void* scratch;       // global scratch of the kModeGlobal kernel
    int64_t scratch_stride;
    int isdmrg;
    int map[3];
};
struct DevSplit {      // an output block computed by several units
    int64_t y_off;     // block offset in the vector
    int64_t part_off;  // first partial copy
    int nparts;
    int size;
};
struct DevPair {       // one (output block, input block) pair
    int64_t x_off;
    int d[3];          // input leg dims in operator order
    int ord[3];        // operator legs in contraction order (ord[2] unused for DMRG)
    int chunk_begin;   // term chunks of this pair
    int chunk_end;
    int w1;            // elements of one first-stage output
    int w2;            // elements of one second-stage output (0 for DMRG)
    int xlds;          // input block staged in LDS
    int wlds;          // stage outputs in LDS (else workgroup global scratch)
    int w1_off;        // first-stage outputs: offset in the LDS pool (after the input block)
    int fuse;          // middle stage fused into stage C (chunks stage alpha O_b, no W2)
};
// Consecutive terms of a pair processed between two barriers: their
// distinct first-stage operators ("runs") and the terms themselves.
// Terms sharing their last operator form a group: stage B sums the group's
// terms into one second-stage output and stage C applies the shared
// operator once per group.  A chunk's terms are ordered by group.  (Terms
// are grouped on the FP32 path; FP64 and DMRG objects have one group per
// term.)
struct DevChunk {
    int term_begin, term_end;
    int run_begin, run_end;  // into the run-lead list; the ncs + 1 group
                             // offsets (terms, chunk-relative) follow at run_end
    int ncs;                 // groups
    int w2_off;              // second-stage outputs: offset from the first-stage outputs
};
struct DevTerm {       // one instruction of one MPO group of a pair
    int64_t op[3];     // element offsets into the three operator arenas
    double alpha;      // instruction prefactor x symmetry prefactor
    int slot;          // first-stage output slot (run index within its chunk)
    int cslot;         // group of its last operator within its chunk
};
struct DevDiagBlock {
    int64_t off;
    int d[3];
    int term_begin;
    int term_end;
    int pad;
};

struct HeffStats {
    int64_t nblocks = 0, npairs = 0, nterms = 0, size = 0;
    double flops = 0;            // reference-order GEMM flops of one matvec
    int64_t op_elems[3] = {0, 0, 0};
    int64_t scratch_per_wg = 0;
    int64_t nchunks = 0;
    int64_t nruns = 0;
    int64_t nunits = 0;
    int64_t nsplit_blocks = 0;
    int64_t pairs_global_x = 0;  // input block read from global memory
    int64_t pairs_global_w = 0;  // stage outputs in global scratch
    int64_t pairs_fused = 0;     // pairs whose middle stage is fused into stage C
    double mfma_steps = 0;       // 16x16x4 matrix-core steps of one matvec
    double stage_steps[3] = {0, 0, 0};
};

// Prices of the cost model that flattens the plan into work lists: they
// decide contraction orders, the fused middle stage and chunking, hence the
// kernels' work.  Defaults measured on MI355X; the autotuner
// (t3ns_heff_autotune) picks per block-size histogram.
struct HeffTuning {
    double chunk_steps = 1000;  // a chunk of terms (barriers, memory latency), in matrix-core steps
    double fuse_over = 2.0;     // fused operand work per fiber tile and k-step, in steps
    int fuse = 1;               // fused middle stage allowed
    int sched = 0;              // unit schedule: 0 cost, 1 block, 2 xcd, 3 xcdblock
    int wpc = 0;                // workgroups per CU (8 / wpc waves each, 1 / wpc of the LDS): 1, 2
                                // or 4; 0 picks by the object's structure
    // The defaults with T3NS_HEFF_CHUNK_STEPS, T3NS_HEFF_FUSE,
    // T3NS_HEFF_FUSE_OVER, T3NS_HEFF_SCHED and T3NS_HEFF_WPC applied.
    static HeffTuning from_env();
};

class HeffDevice;

// Matvecs of several objects of one precision in one launch per kernel
// (cross-solve batching): their work units share the persistent grid, so
// small objects fill the GPU together.  Each object's result is bitwise the
// one of its own matvec (a unit's work does not depend on its neighbours).
// The buffers are reused by the next call; calls on one batch must share a
// stream.
class HeffBatch {
public:
    void matvec(HeffDevice* const* hs, const double* const* x, double* const* y, int n,
                hipStream_t s);

private:
    template <typename T, typename OT>
    void launch(HeffDevice* const* hs, const T* const* x, T* const* y, int n, hipStream_t s);
    template <typename T, typename OT, int WPC>
    void launch_grid(const std::vector<DevBlock> (&units)[2], int base, int shape, hipStream_t s);
    template <typename T, typename OT, int WPC>
    void launch_glob(HeffDevice& h, const T* x, T* y, hipStream_t s);
    int ncu_ = 256;
    DevBuf<DevBlock> units_;
    DevBuf<HeffRef> refs_;
    DevBuf<int> counter_;
    DevBuf<int> seg_;
};

class HeffDevice {
    friend class HeffBatch;

public:
    // Builds the plan tables and uploads the operator arenas (host data in
    // in.ops[k]->tel).  `stream` may be null (default stream); `tuning`
    // null: HeffTuning::from_env().
    HeffDevice(const HeffInput& in, int precision, hipStream_t stream,
               const HeffTuning* tuning = nullptr);
    // An FP32 object on the plan and FP64 operator arenas of an FP64 one
    // built from the same `in` (the arenas copied and converted on the
    // device, the plan copied): only its own work lists are built on the
    // host, and no operator data crosses the host link.
    HeffDevice(const HeffDevice& base, const HeffInput& in, int precision, hipStream_t stream);
    ~HeffDevice();
    HeffDevice(const HeffDevice&) = delete;
    HeffDevice& operator=(const HeffDevice&) = delete;

    int64_t size() const { return size_; }
    const HeffStats& stats() const { return stats_; }
    const HeffPlan& plan() const { return plan_; }
    int precision() const { return prec_; }

    // Device pointers, FP64 in and out.
    void matvec(const double* dvec, double* dres, hipStream_t s);
    // Replays matvecs from captured HIP graphs, one per (input, output)
    // pair of arrays (up to kMaxGraphs, oldest dropped first); off by
    // default.
    void set_graphs(bool on);
    bool graphs() const { return use_graphs_; }
    static constexpr size_t kMaxGraphs = 64;
    void diagonal(double* ddiag, hipStream_t s);
    // Rebuilds the work lists under another tuning (`in` as given to the
    // constructor; operator arenas are kept).  Results change by rounding
    // only (contraction orders may differ).
    void retune(const HeffInput& in, const HeffTuning& t, hipStream_t s);
    const HeffTuning& tuning() const { return tune_; }
    int wpc() const { return wpc_; }  // grid shape in use (tuning().wpc, or the one picked)

private:
    void build_lists(const HeffInput& in, hipStream_t stream);
    int build_lists_at(const HeffInput& in, hipStream_t stream, int wpc, bool pick);
    void enqueue(const double* dvec, double* dres, hipStream_t s);
    void drop_graphs();
    template <typename T, typename OT, int WPC>
    void launch(const T* x, T* y, T* part, hipStream_t s);
    template <typename T, typename OT>
    void launch_any(const T* x, T* y, T* part, hipStream_t s);
    void build_ref();

    HeffPlan plan_;
    HeffStats stats_;
    HeffTuning tune_;
    HeffRef href_{};        // the kernels' view of this object
    std::vector<DevBlock> hunits_;  // the unit list (host copy) and unit costs
    std::vector<double> hucost_;
    std::vector<std::vector<std::vector<float>>> bscale_;  // BF16 block scales [leg][op][block]
    int prec_;
    int isdmrg_;
    int map_[3];
    int64_t size_;
    int nwg_glob_ = 1;
    int ncu_ = 256;
    int wpc_ = 1;     // workgroups per CU of the lists built (HeffTuning::wpc)
    int nfast_ = 0;   // units whose pairs all fit in LDS (first in the unit list)
    int nxg_ = 0;     // units whose stage outputs fit in LDS but input blocks do not
    int nglob_ = 0;   // units that use global scratch
    DevBuf<DevBlock> blocks_;
    DevBuf<DevPair> pairs_;
    DevBuf<DevChunk> chunks_;
    DevBuf<int> runs_;
    DevBuf<DevTerm> terms_;
    DevBuf<DevDiagBlock> dblocks_;
    DevBuf<DevTerm> dterms_;
    DevBuf<double> ops64_[3];
    DevBuf<float> ops32_[3];
    DevBuf<uint16_t> ops16_[3];  // BF16 arenas, each block divided by its power-of-two scale
    DevBuf<double> scratch64_;
    DevBuf<float> scratch32_;
    DevBuf<float> x32_, y32_;
    // Unit schedule: per kernel, up to kSegments segments of the unit list
    // (one per XCD), each with its own counter.
    static constexpr int kSegments = 8;
    DevBuf<int> counter_;  // next unit of each segment
    DevBuf<int> seg_;      // segment boundaries, kSegments + 1 per kernel
    int nseg_[3] = {1, 1, 1};
    DevBuf<DevSplit> splits_;
    DevBuf<double> part64_;
    DevBuf<float> part32_;
    int64_t part_elems_ = 0;
    int nsplits_ = 0;
    int ndblocks_ = 0;
    struct Graph {
        const double* x;
        double* y;
        hipGraphExec_t exec;
    };
    bool use_graphs_ = false;
    std::vector<Graph> graphs_;
};

}  // namespace t3
