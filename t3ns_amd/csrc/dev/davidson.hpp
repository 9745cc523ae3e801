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
    DeviceDavidson(int64_t size, int max_vecs, int keep_deflate, hipStream_t s);
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
    void combine(const double* A, int n, const double* coef_host, double* out, double beta);
    double norm(const double* v);
    void orthonormalise(double* t);
    void expand(int m);
    int eigsolve();
    void deflate();
    double ritz_and_correct(double* x, const double* diag);
    DavidsonResult iterate(double* x, const double* diag, const DevMatvec& mv, double tol,
                           int max_its, double rel_floor, int its, double residual);

    int64_t size_;
    int mv_, keep_;
    hipStream_t s_;
    DevBuf<double> V_, VA_, t_, tmp_;
    DevBuf<double> coef_, partial_, out_;
    std::vector<double> sub_, eigv_, eigval_;
    int m_ = 0;
    int nblk_ = 0;
};

// Symmetric eigendecomposition of a small dense matrix (column-major n x n
// with leading dimension ld) by cyclic Jacobi rotations; eigenvalues in
// ascending order, eigenvectors as columns of V (ld).  Returns 0 on
// success.
int jacobi_eigh(int n, const double* A, int ld, double* w, double* V);

}  // namespace t3
