// Shared host-side helpers of the femahoy_amd library: HIP error checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace fem {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define FEM_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::fem::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                               ":" + std::to_string(__LINE__) + " in " #call);                         \
    } while (0)

// An owned device allocation.
// Device bytes of the Buffers allocated on this thread while a meter is set: live bytes and their
// peak (the footprint of a solver's construction).
struct AllocMeter {
    int64_t live = 0, peak = 0;
};
inline thread_local AllocMeter* alloc_meter = nullptr;

class Buffer {
   public:
    Buffer() = default;
    explicit Buffer(size_t bytes) { alloc(bytes); }
    ~Buffer() { release(); }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes) FEM_HIP(hipMalloc(&p_, bytes));
        n_ = bytes;
        if (alloc_meter) {
            alloc_meter->live += (int64_t)bytes;
            if (alloc_meter->live > alloc_meter->peak) alloc_meter->peak = alloc_meter->live;
        }
    }
    void release() {
        if (p_) (void)hipFree(p_);
        if (alloc_meter) alloc_meter->live -= (int64_t)n_;
        p_ = nullptr;
        n_ = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p_);
    }
    void* get() const { return p_; }
    size_t bytes() const { return n_; }

   private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// A typed device array.
template <class T>
class DArray {
   public:
    DArray() = default;
    explicit DArray(size_t n) { resize(n); }
    void resize(size_t n) {
        buf_.alloc(n * sizeof(T));
        n_ = n;
    }
    T* data() const { return buf_.as<T>(); }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }
    void upload(const T* host, size_t n, hipStream_t s = nullptr) {
        if (n > n_) throw Error("upload beyond the array");
        FEM_HIP(hipMemcpyAsync(data(), host, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* host, size_t n, hipStream_t s = nullptr) const {
        if (n > n_) throw Error("download beyond the array");
        FEM_HIP(hipMemcpyAsync(host, data(), n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

   private:
    Buffer buf_;
    size_t n_ = 0;
};

}  // namespace fem

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace fem {

// Rust's `{:e}` of an f64 (the reference's error texts, e.g. crates/fem/src/cholesky.rs:94-98):
// the shortest digits that read back to the same value, in scientific notation with the exponent
// written without sign padding or leading zeros ("1e0", "2.5e-3", "-4.25e12"; "NaN", "inf").
inline std::string rust_e(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char num[64];
    for (int prec = 0; prec <= 17; ++prec) {
        std::snprintf(num, sizeof num, "%.*e", prec, v);
        if (std::strtod(num, nullptr) == v) break;
    }
    std::string t(num);
    const size_t e = t.find('e');
    if (e == std::string::npos) return t;
    return t.substr(0, e) + "e" + std::to_string(std::atoi(t.c_str() + e + 1));
}

}  // namespace fem
