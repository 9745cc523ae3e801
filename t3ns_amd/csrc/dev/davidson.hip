// Device-resident Davidson (src/davidson.c restated for HBM-resident
// vectors; see davidson.hpp).
#include "davidson.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <mutex>
#include <stdexcept>
#include <string>

namespace t3 {

namespace {
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;
constexpr double kDiagCutoff = 1e-12;  // src/davidson.c:33

// partial[j * nb + b] = sum over block b's grid-stride share of
// A[j*lda + i] * x[i], j < n.
__global__ void __launch_bounds__(kRedThreads)
    dots_kernel(const double* __restrict__ A, int64_t lda, int n, const double* __restrict__ x,
                int64_t size, double* __restrict__ partial) {
    __shared__ double red[kRedThreads];
    for (int j = 0; j < n; ++j) {
        const double* a = A + (int64_t)j * lda;
        double s = 0;
        for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < size;
             i += (int64_t)gridDim.x * kRedThreads)
            s += a[i] * x[i];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// out[j] = sum_b partial[j * nb + b], in block order.
__global__ void __launch_bounds__(kRedThreads)
    finish_kernel(const double* __restrict__ partial, int nb, int n, double* __restrict__ out) {
    __shared__ double red[kRedThreads];
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += kRedThreads) s += partial[(int64_t)j * nb + b];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = kRedThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = red[0];
        __syncthreads();
    }
}

// out[i] = beta * out[i] + sum_j (cscale coef[j]) A[j*size + i]
__global__ void combine_kernel(const double* __restrict__ A, int64_t size, int n,
                               const double* __restrict__ coef, double* __restrict__ out,
                               double beta, double cscale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < size;
         i += (int64_t)gridDim.x * blockDim.x) {
        double s = beta == 0.0 ? 0.0 : beta * out[i];
        for (int j = 0; j < n; ++j) s += (cscale * coef[j]) * A[(int64_t)j * size + i];
        out[i] = s;
    }
}

// v *= 1 / sqrt(*sq) (v /= its norm, the squared norm on the device); a
// vanished vector stays zero (the caller checks *sq)
__global__ void unit_kernel(double* __restrict__ v, int64_t n, const double* __restrict__ sq) {
    const double q = *sq;
    const double a = q > 0 ? 1.0 / sqrt(q) : 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v[i] *= a;
}

// t = y - theta x
__global__ void residual_kernel(const double* __restrict__ y, const double* __restrict__ x,
                                double theta, double* __restrict__ t, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        t[i] = y[i] - theta * x[i];
}

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__device__ __forceinline__ double precond_inv(double diff) {
    // 1/diff with the cutoff of src/davidson.c:230-236
    return fabs(diff) > kDiagCutoff ? 1.0 / diff : (diff < 0 ? -1.0 : 1.0) / kDiagCutoff;
}

// uK = r / (D - theta); partial sums of uK*t and uK*r (src/davidson.c:223-242)
__global__ void __launch_bounds__(kRedThreads)
    precond_dots_kernel(const double* __restrict__ r, const double* __restrict__ t,
                        const double* __restrict__ diag, double theta, int64_t size,
                        double* __restrict__ partial) {
    __shared__ double red[2][kRedThreads];
    double s0 = 0, s1 = 0;
    for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < size;
         i += (int64_t)gridDim.x * kRedThreads) {
        const double uK = r[i] * precond_inv(diag[i] - theta);
        s0 += uK * t[i];
        s1 += uK * r[i];
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = red[0][0];
        partial[gridDim.x + blockIdx.x] = red[1][0];
    }
}

// t = -(t + alpha r) / (D - theta), alpha = -uk[0] / uk[1]
// (src/davidson.c:243-256)
__global__ void precond_apply_kernel(double* __restrict__ t, const double* __restrict__ r,
                                     const double* __restrict__ diag, double theta,
                                     const double* __restrict__ uk, int64_t size) {
    const double alpha = -uk[0] / uk[1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < size;
         i += (int64_t)gridDim.x * blockDim.x)
        t[i] = -(t[i] + alpha * r[i]) * precond_inv(diag[i] - theta);
}

int grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }
}  // namespace

int jacobi_eigh(int n, const double* A, int ld, double* w, double* V, double* wa, double* wv) {
    if (n < 1 || n > kMaxDavidsonVecs) return 1;
    double* a = wa;
    double* v = wv;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            a[i + (size_t)j * n] = A[std::min(i, j) + (size_t)std::max(i, j) * ld];
            v[i + (size_t)j * n] = i == j ? 1.0 : 0.0;
        }
    auto at = [&](int i, int j) -> double& { return a[i + (size_t)j * n]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                tot += at(i, j) * at(i, j);
                if (i != j) off += at(i, j) * at(i, j);
            }
        if (off <= 1e-30 * std::max(tot, 1e-300)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = at(p, q);
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (at(q, q) - at(p, p)) / (2 * apq);
                const double tt = (theta >= 0 ? 1.0 : -1.0) /
                                  (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(tt * tt + 1), s = tt * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = at(k, p), akq = at(k, q);
                    at(k, p) = c * akp - s * akq;
                    at(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = at(p, k), aqk = at(q, k);
                    at(p, k) = c * apk - s * aqk;
                    at(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[k + (size_t)p * n], vkq = v[k + (size_t)q * n];
                    v[k + (size_t)p * n] = c * vkp - s * vkq;
                    v[k + (size_t)q * n] = s * vkp + c * vkq;
                }
            }
    }
    // ascending eigenvalues, equal ones in index order (stable insertion sort)
    int idx[kMaxDavidsonVecs];
    for (int j = 0; j < n; ++j) {
        int k = j;
        while (k > 0 && at(idx[k - 1], idx[k - 1]) > at(j, j)) {
            idx[k] = idx[k - 1];
            --k;
        }
        idx[k] = j;
    }
    for (int j = 0; j < n; ++j) {
        w[j] = at(idx[j], idx[j]);
        for (int i = 0; i < n; ++i) V[i + (size_t)j * ld] = v[i + (size_t)idx[j] * n];
    }
    for (int j = 0; j < n; ++j)
        if (!std::isfinite(w[j])) return 1;
    return 0;
}

namespace {
constexpr size_t kAlign = 32;  // doubles: 256 bytes
size_t rounded(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
}  // namespace

void DeviceDavidson::setup(int64_t size, int max_vecs, int keep_deflate) {
    if (size <= 0 || max_vecs < 2 || keep_deflate < 1 || keep_deflate >= max_vecs)
        throw std::invalid_argument("davidson: bad sizes");
    if (max_vecs > kMaxDavidsonVecs)
        throw std::invalid_argument("davidson: max_vecs above " + std::to_string(kMaxDavidsonVecs));
    size_ = size;
    mv_ = max_vecs;
    keep_ = keep_deflate;
    if (max_vecs > size) mv_ = (int)std::max<int64_t>(2, size);
    if (keep_ >= mv_) keep_ = mv_ - 1;
    nblk_ = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (size + kRedThreads - 1) / kRedThreads));
    std::fill(sub_, sub_ + kMaxDavidsonVecs * kMaxDavidsonVecs, 0.0);
    std::fill(eigv_, eigv_ + kMaxDavidsonVecs * kMaxDavidsonVecs, 0.0);
    std::fill(eigval_, eigval_ + kMaxDavidsonVecs, 0.0);
}

size_t DeviceDavidson::workspace_elems(int64_t size, int max_vecs, int keep_deflate) {
    int mv = max_vecs, keep = keep_deflate;
    if (mv > size) mv = (int)std::max<int64_t>(2, size);
    if (keep >= mv) keep = mv - 1;
    const size_t nb = (size_t)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (size + kRedThreads - 1) / kRedThreads));
    return 2 * rounded((size_t)size * mv) + rounded(size) + rounded((size_t)size * keep) +
           rounded(nb * (mv + 2)) + rounded(mv) + rounded(mv + 2) + rounded(kScalars);
}

void DeviceDavidson::carve(double* ws) {
    double* p = ws;
    auto take = [&](size_t n) {
        double* q = p;
        p += rounded(n);
        return q;
    };
    V_ = take((size_t)size_ * mv_);
    VA_ = take((size_t)size_ * mv_);
    t_ = take(size_);
    tmp_ = take((size_t)size_ * keep_);
    partial_ = take((size_t)nblk_ * (mv_ + 2));
    coef_ = take(mv_);
    out_ = take(mv_ + 2);
    scal_ = take(kScalars);
}

DeviceDavidson::DeviceDavidson(int64_t size, int max_vecs, int keep_deflate, hipStream_t s)
    : s_(s) {
    setup(size, max_vecs, keep_deflate);
    own_.alloc(workspace_elems(size, max_vecs, keep_deflate));
    carve(own_.get());
}

DeviceDavidson::DeviceDavidson(int64_t size, int max_vecs, int keep_deflate, hipStream_t s,
                               double* ws, size_t ws_elems)
    : s_(s) {
    setup(size, max_vecs, keep_deflate);
    if (!ws || ws_elems < workspace_elems(size, max_vecs, keep_deflate))
        throw std::invalid_argument("davidson: workspace too small");
    carve(ws);
}

void DeviceDavidson::dots(const double* A, int64_t lda, int n, const double* b, double* host_out) {
    if (n == 0) return;
    hipLaunchKernelGGL(dots_kernel, dim3(nblk_), dim3(kRedThreads), 0, s_, A, lda, n, b, size_,
                       partial_);
    hipLaunchKernelGGL(finish_kernel, dim3(std::min(n, 64)), dim3(kRedThreads), 0, s_,
                       partial_, nblk_, n, out_);
    T3_HIP_CHECK(hipMemcpyAsync(host_out, out_, n * sizeof(double), hipMemcpyDeviceToHost, s_));
    T3_HIP_CHECK(hipStreamSynchronize(s_));
}

void DeviceDavidson::combine(const double* A, int n, const double* coef_host, double* out,
                             double beta) {
    T3_HIP_CHECK(hipMemcpyAsync(coef_, coef_host, n * sizeof(double), hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(combine_kernel, dim3(grid(size_)), dim3(256), 0, s_, A, size_, n,
                       coef_, out, beta, 1.0);
}

// dst[j] = A[:, j] . b (j < n) into device memory: no host round trip.
void DeviceDavidson::dots_dev(const double* A, int64_t lda, int n, const double* b, double* dst) {
    hipLaunchKernelGGL(dots_kernel, dim3(nblk_), dim3(kRedThreads), 0, s_, A, lda, n, b, size_,
                       partial_);
    hipLaunchKernelGGL(finish_kernel, dim3(std::min(n, 64)), dim3(kRedThreads), 0, s_, partial_,
                       nblk_, n, dst);
}

double DeviceDavidson::norm(const double* v) {
    double h;
    dots(v, size_, 1, v, &h);
    return std::sqrt(h);
}

// t <- orthonormal complement of t against V[:, :m] (two CGS passes), on
// the device without a host round trip: the coefficients stay in device
// memory, and t's squared norm before scaling is left in scal_[0] for the
// caller to check (a vanished vector stays zero).
void DeviceDavidson::orthonormalise(double* t) {
    for (int pass = 0; pass < 2 && m_ > 0; ++pass) {
        dots_dev(V_, size_, m_, t, coef_);
        hipLaunchKernelGGL(combine_kernel, dim3(grid(size_)), dim3(256), 0, s_, V_, size_, m_,
                           coef_, t, 1.0, -1.0);
    }
    dots_dev(t, size_, 1, t, scal_);
    hipLaunchKernelGGL(unit_kernel, dim3(grid(size_)), dim3(256), 0, s_, t, size_, scal_);
}

// Column m of the projected matrix: V[:, i] . VA[:, m], i <= m
// (src/davidson.c:132-145).  With `sq`, the squared norm orthonormalise()
// left on the device comes back in the same host round trip.
void DeviceDavidson::expand(int m, double* sq) {
    double* c = hc_;
    dots_dev(V_, size_, m + 1, VA_ + (int64_t)m * size_, out_);
    T3_HIP_CHECK(hipMemcpyAsync(c, out_, (m + 1) * sizeof(double), hipMemcpyDeviceToHost, s_));
    if (sq) T3_HIP_CHECK(hipMemcpyAsync(sq, scal_, sizeof(double), hipMemcpyDeviceToHost, s_));
    T3_HIP_CHECK(hipStreamSynchronize(s_));
    for (int i = 0; i <= m; ++i) sub_[i + (size_t)m * mv_] = c[i];
}

int DeviceDavidson::eigsolve() {
    return jacobi_eigh(m_, sub_, mv_, eigval_, eigv_, ja_, jv_);
}

// Keep the keep_ lowest Ritz vectors (src/davidson.c:164-185).
void DeviceDavidson::deflate() {
    for (int which = 0; which < 2; ++which) {
        double* B = which == 0 ? V_ : VA_;
        for (int k = 0; k < keep_; ++k)
            combine(B, m_, eigv_ + (size_t)k * mv_, tmp_ + (int64_t)k * size_, 0.0);
        hipLaunchKernelGGL(copy_kernel, dim3(grid(size_ * keep_)), dim3(256), 0, s_, tmp_, B,
                           size_ * keep_);
    }
    std::fill(sub_, sub_ + (size_t)mv_ * mv_, 0.0);
    m_ = 0;
    while (m_ < keep_) {
        expand(m_);
        ++m_;
    }
}

std::pair<double, double> DeviceDavidson::rayleigh(const double* x, const DevMatvec& mv) {
    double* y = VA_;
    mv(x, y, s_);
    double xx, xy;
    dots(x, size_, 1, x, &xx);
    dots(x, size_, 1, y, &xy);
    if (!(xx > 0)) throw std::runtime_error("davidson: zero vector");
    const double theta = xy / xx;
    hipLaunchKernelGGL(residual_kernel, dim3(grid(size_)), dim3(256), 0, s_, y, x, theta, t_,
                       size_);
    return {theta, norm(t_) / std::sqrt(xx)};
}

double DeviceDavidson::relative_difference(const double* x, const DevMatvec& a, const DevMatvec& b) {
    double* ya = VA_;
    double* yb = t_;
    a(x, ya, s_);
    b(x, yb, s_);
    const double na = norm(ya);
    hipLaunchKernelGGL(residual_kernel, dim3(grid(size_)), dim3(256), 0, s_, ya, yb, 1.0, tmp_,
                       size_);
    return na > 0 ? norm(tmp_) / na : 0.0;
}

// Ritz vector x = V y, residual t = VA y - theta x (src/davidson.c:187-202),
// then the preconditioned correction of t against x (src/davidson.c:
// 204-257).  Returns the residual norm; theta is eigval_[0].
//  One host round trip: the Ritz coefficients go down once, the residual's
// squared norm comes back; the preconditioner's two dot products stay on
// the device.
double DeviceDavidson::ritz_and_correct(double* x, const double* diag) {
    const double theta = eigval_[0];
    const dim3 g(grid(size_)), b(256);
    T3_HIP_CHECK(hipMemcpyAsync(coef_, eigv_, m_ * sizeof(double), hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(combine_kernel, g, b, 0, s_, V_, size_, m_, coef_, x, 0.0, 1.0);
    hipLaunchKernelGGL(combine_kernel, g, b, 0, s_, VA_, size_, m_, coef_, t_, 0.0, 1.0);
    hipLaunchKernelGGL(combine_kernel, g, b, 0, s_, V_, size_, m_, coef_, t_, 1.0, -theta);
    dots_dev(t_, size_, 1, t_, scal_ + 1);
    hipLaunchKernelGGL(precond_dots_kernel, dim3(nblk_), dim3(kRedThreads), 0, s_, x, t_,
                       diag, theta, size_, partial_);
    hipLaunchKernelGGL(finish_kernel, dim3(2), dim3(kRedThreads), 0, s_, partial_, nblk_,
                       2, scal_ + 2);
    double r2;
    T3_HIP_CHECK(hipMemcpyAsync(&r2, scal_ + 1, sizeof(double), hipMemcpyDeviceToHost, s_));
    T3_HIP_CHECK(hipStreamSynchronize(s_));
    hipLaunchKernelGGL(precond_apply_kernel, g, b, 0, s_, t_, x, diag, theta, scal_ + 2, size_);
    return std::sqrt(r2);
}

DavidsonResult DeviceDavidson::iterate(double* x, const double* diag, const DevMatvec& mv,
                                       double tol, int max_its, double rel_floor, int its,
                                       double residual) {
    DavidsonResult R;
    bool floor_hit = false;
    while (residual > tol && its < max_its) {
        if (rel_floor > 0 && m_ > 0 && residual <= rel_floor * std::fabs(eigval_[0])) {
            floor_hit = true;
            break;
        }
        orthonormalise(t_);
        double* Vm = V_ + (int64_t)m_ * size_;
        double* VAm = VA_ + (int64_t)m_ * size_;
        hipLaunchKernelGGL(copy_kernel, dim3(grid(size_)), dim3(256), 0, s_, t_, Vm, size_);
        mv(Vm, VAm, s_);
        double sq = 0;
        expand(m_, &sq);
        if (!(sq > 0)) throw std::runtime_error("davidson: search vector vanished");
        ++m_;
        if (eigsolve() != 0) {
            R.status = -1;
            return R;
        }
        if (m_ == mv_) {
            deflate();
            if (eigsolve() != 0) {
                R.status = -1;
                return R;
            }
        }
        residual = ritz_and_correct(x, diag);
        ++its;
    }
    T3_HIP_CHECK(hipStreamSynchronize(s_));
    R.energy = m_ > 0 ? eigval_[0] : 0.0;
    R.residual = residual;
    R.its = its;
    R.status = residual <= tol ? 0 : floor_hit ? 2 : 1;
    return R;
}

DavidsonResult DeviceDavidson::solve(double* x, const double* diag, const DevMatvec& mv,
                                     double tol, int max_its, double rel_floor) {
    m_ = 0;
    hipLaunchKernelGGL(copy_kernel, dim3(grid(size_)), dim3(256), 0, s_, x, t_, size_);
    return iterate(x, diag, mv, tol, max_its, rel_floor, 0, HUGE_VAL);
}

DavidsonResult DeviceDavidson::resume(double* x, const double* diag, const DevMatvec& mv,
                                      double tol, int max_its) {
    if (m_ == 0) return solve(x, diag, mv, tol, max_its);
    if (m_ > keep_) {
        // keep the lowest Ritz vectors of the current projected problem
        for (int k = 0; k < keep_; ++k)
            combine(V_, m_, eigv_ + (size_t)k * mv_, tmp_ + (int64_t)k * size_, 0.0);
        hipLaunchKernelGGL(copy_kernel, dim3(grid(size_ * keep_)), dim3(256), 0, s_, tmp_,
                           V_, size_ * keep_);
        m_ = keep_;
    }
    std::fill(sub_, sub_ + (size_t)mv_ * mv_, 0.0);
    for (int k = 0; k < m_; ++k) {
        mv(V_ + (int64_t)k * size_, VA_ + (int64_t)k * size_, s_);
        expand(k);
    }
    DavidsonResult R;
    if (eigsolve() != 0) {
        R.status = -1;
        return R;
    }
    const double residual = ritz_and_correct(x, diag);
    return iterate(x, diag, mv, tol, max_its, 0.0, 1, residual);
}

// ---------------------------------------------------------------------------
// Davidson without allocation.  The solver's device buffers live in this
// file-scope array: every device the library runs on holds its own copy
// (kNoAllocBytes) until the process ends.  Host code may not use the
// array's name as a device address; each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that pointer is handed to kernels and copies.  Every call on a
// device shares that device's array and leaves its contents behind: the
// solver writes every buffer before reading it, so nothing needs clearing,
// and calls that overlap on one device race.
namespace {
__device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* scratch_for_current_device() {
    int dev = 0;
    T3_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("HIP error: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        T3_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

DavidsonResult davidson_noalloc(int64_t size, double* dx, const double* ddiag,
                                const DevMatvec& mv, hipStream_t s, double tol, int max_vecs,
                                int keep_deflate, int max_its) {
    const size_t cap = kNoAllocBytes / sizeof(double);
    if (size <= 0 || max_vecs < 2 || keep_deflate < 1 || keep_deflate >= max_vecs)
        throw std::invalid_argument("davidson: bad sizes");
    const size_t need = DeviceDavidson::workspace_elems(size, max_vecs, keep_deflate);
    if (need > cap)
        throw std::invalid_argument(
            "davidson_noalloc: " + std::to_string(size) + " elements x " + std::to_string(max_vecs) +
            " vectors need " + std::to_string(need * sizeof(double)) +
            " bytes of workspace, more than the fixed " + std::to_string(kNoAllocBytes) + " bytes");
    double* ws = scratch_for_current_device();
    DeviceDavidson dav(size, max_vecs, keep_deflate, s, ws, cap);
    return dav.solve(dx, ddiag, mv, tol, max_its);
}

}  // namespace t3
