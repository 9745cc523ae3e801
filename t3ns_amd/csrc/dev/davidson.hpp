// Device-resident Davidson eigensolver for the lowest eigenpair of a
// symmetric operator given as a device matvec.  Follows the algorithm of
// src/davidson.c (Arbenz, lecture notes ch. 12, algorithm 12.1): search
// vectors V and their images VA in HBM, the projected matrix and its
// eigenproblem on the host (m <= max_vecs), deflation to keep_deflate
// vectors when the basis is full, the diagonal (Olsen) preconditioner of
// src/davidson.c:223-257 with its 1e-12 cutoff.  Orthogonalisation is
// classical Gram-Schmidt applied twice (batched dot products and one
// combination per pass) instead of the reference's sequential modified
// Gram-Schmidt; both give an orthonormal basis to working precision.
// All reductions are deterministic (fixed grid, fixed summation order).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

#include "hip_util.hpp"

namespace t3 {

// Largest search space (max_vecs) the solver takes; its host-side
// projected problem lives in fixed arrays of this size (no heap use).
constexpr int kMaxDavidsonVecs = 64;

// y = A x on device pointers (FP64), on the given stream.
using DevMatvec = std::function<void(const double* x, double* y, hipStream_t s)>;

struct DavidsonResult {
    double energy = 0;
    double residual = 0;
    int its = 0;
    int status = 0;  // 0 converged, 1 max_its reached, 2 at the precision floor, -1 failure
};

class DeviceDavidson {
public:
    // Allocates its device buffers (hipMalloc).
    DeviceDavidson(int64_t size, int max_vecs, int keep_deflate, hipStream_t s);
    // Takes its device buffers from `ws` (ws_elems doubles, at least
    // workspace_elems(size, max_vecs, keep_deflate)) and allocates nothing.
    DeviceDavidson(int64_t size, int max_vecs, int keep_deflate, hipStream_t s, double* ws,
                   size_t ws_elems);
    // Device doubles the solver needs for these sizes (each buffer rounded
    // up to 256 bytes).
    static size_t workspace_elems(int64_t size, int max_vecs, int keep_deflate);
    // x (device, size): initial guess in, eigenvector out.  diag (device):
    // the operator's diagonal for the preconditioner.
    // rel_floor > 0: also stop (status 2) once the residual is below
    // rel_floor * |theta|, the accuracy limit of a reduced-precision matvec
    // whose relative error is about rel_floor.
    DavidsonResult solve(double* x, const double* diag, const DevMatvec& mv, double tol,
                         int max_its, double rel_floor = 0);
    // Continues the last solve with another matvec: the basis is deflated
    // to its keep_deflate lowest Ritz vectors, their images are recomputed
    // with mv, and the iteration goes on from there.
    DavidsonResult resume(double* x, const double* diag, const DevMatvec& mv, double tol,
                          int max_its);
    // Rayleigh quotient of x (device) under mv and the norm of its
    // residual, both relative to |x|: (x.Ax / x.x, |Ax - theta x| / |x|).
    std::pair<double, double> rayleigh(const double* x, const DevMatvec& mv);
    // |A x - B x| / |A x| for two matvecs of one operator.
    double relative_difference(const double* x, const DevMatvec& a, const DevMatvec& b);

private:
    void dots(const double* A, int64_t lda, int n, const double* b, double* host_out);
    void dots_dev(const double* A, int64_t lda, int n, const double* b, double* dst);
    void combine(const double* A, int n, const double* coef_host, double* out, double beta);
    double norm(const double* v);
    void orthonormalise(double* t);
    void expand(int m, double* sq = nullptr);
    int eigsolve();
    void deflate();
    double ritz_and_correct(double* x, const double* diag);
    DavidsonResult iterate(double* x, const double* diag, const DevMatvec& mv, double tol,
                           int max_its, double rel_floor, int its, double residual);

    void setup(int64_t size, int max_vecs, int keep_deflate);
    void carve(double* ws);

    int64_t size_;
    int mv_, keep_;
    hipStream_t s_;
    DevBuf<double> own_;  // the device buffers when the solver allocates them
    double *V_ = nullptr, *VA_ = nullptr, *t_ = nullptr, *tmp_ = nullptr;
    double *coef_ = nullptr, *partial_ = nullptr, *out_ = nullptr;
    // device scalars: |t|^2 before normalising, |r|^2, the preconditioner's
    // two dot products
    static constexpr int kScalars = 8;
    double* scal_ = nullptr;
    // projected matrix, its eigenvectors (both mv_ x mv_, leading dimension
    // mv_), eigenvalues, host coefficients and the eigensolver's work
    double sub_[kMaxDavidsonVecs * kMaxDavidsonVecs];
    double eigv_[kMaxDavidsonVecs * kMaxDavidsonVecs];
    double eigval_[kMaxDavidsonVecs];
    double hc_[kMaxDavidsonVecs + 2];
    double ja_[kMaxDavidsonVecs * kMaxDavidsonVecs];
    double jv_[kMaxDavidsonVecs * kMaxDavidsonVecs];
    int m_ = 0;
    int nblk_ = 0;
};

// Symmetric eigendecomposition of a small dense matrix (column-major n x n
// with leading dimension ld, n <= kMaxDavidsonVecs) by cyclic Jacobi
// rotations; eigenvalues in ascending order, eigenvectors as columns of V
// (ld).  wa, wv: work arrays of n * n doubles.  Returns 0 on success.
int jacobi_eigh(int n, const double* A, int ld, double* w, double* V, double* wa, double* wv);

// Davidson on caller data without any allocation: the solver's device
// buffers come from a file-scope device array of kNoAllocBytes on the
// current device (shared by every such call on that device).  Throws
// std::invalid_argument when the sizes need more than that.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
DavidsonResult davidson_noalloc(int64_t size, double* dx, const double* ddiag,
                                const DevMatvec& mv, hipStream_t s, double tol, int max_vecs,
                                int keep_deflate, int max_its);

}  // namespace t3
