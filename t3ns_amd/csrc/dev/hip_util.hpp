// HIP runtime helpers: error checking and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace t3 {

#define T3_HIP_CHECK(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) +    \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__));   \
    } while (0)

template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t n) {
        release();
        if (n) T3_HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
        n_ = n;
    }
    // Grows (never shrinks) the allocation; contents are not preserved.
    void reserve(size_t n) {
        if (n > n_) alloc(n);
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T* h, size_t n, hipStream_t s = nullptr) {
        reserve(n);
        if (n) T3_HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s = nullptr) { upload(v.data(), v.size(), s); }
    // A device copy of `o` (same size), enqueued on s.
    void copy_from(const DevBuf& o, hipStream_t s = nullptr) {
        alloc(o.n_);
        if (o.n_) T3_HIP_CHECK(hipMemcpyAsync(p_, o.p_, o.n_ * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace t3
