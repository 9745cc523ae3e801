// Device execution of the SVD split (host/decompose.hpp): gather of the
// sector matrices, one-sided Jacobi SVD of every sector on the device,
// truncation on the host, scatter into the two tensors.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../host/decompose.hpp"
#include "hip_util.hpp"
#include "renorm_dev.hpp"

namespace t3 {

struct SplitResult {
    // U: the remaining sites (singular values multiplied in, normalised);
    // V: the split site.  Quantum numbers refer to the cut bond's kept
    // sectors (sectors without a kept value removed).
    SiteTensor U, V;
    Symsecs bond;                        // the cut bond after truncation
    std::vector<std::vector<double>> sing;  // per original sector, all singular values (descending)
    std::vector<int> kept;               // per original sector
    SelectRes res;
    double norm_U = 0;                   // norm removed from U
    int sweeps = 0;                      // Jacobi sweeps until convergence
};

SplitResult run_split(const SplitInput& in, const SplitPlan& plan, const std::vector<double>& A,
                      hipStream_t s, RenormWork* work);

struct QRResult {
    // Q: the one-site tensor with the bond's dims reduced to min(rows, dim)
    // per sector; B: the centre with R multiplied in.  Same blocks and
    // quantum numbers as the inputs (blocks of empty sectors have size 0).
    SiteTensor Q, B;
    std::vector<int> dims;                  // new dims of the bond per sector
    std::vector<std::vector<double>> sing;  // singular values of R per sector
    double entropy = 0;                     // von Neumann entropy of the R spectrum
};
QRResult run_qr(const QRInput& in, const QRPlan& plan, hipStream_t s, RenormWork* work);

// One-sided Jacobi SVD of column-major matrices on the device (rows >=
// columns per matrix).  B is overwritten by U * Sigma (columns unsorted),
// V (n x n) receives the right singular vectors.  Returns the sweeps run.
struct JacobiMat {
    int64_t b, v;  // offsets of B (m x n) and V (n x n) in the arena
    int m, n;
};
int jacobi_svd(double* arena, const std::vector<JacobiMat>& mats, hipStream_t s, int max_sweeps,
               double tol);

}  // namespace t3
