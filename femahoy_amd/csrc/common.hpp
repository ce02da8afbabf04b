// Shared host-side helpers of the femahoy_amd library: HIP error checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

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

// One device allocation made up front and handed out in pieces: the serving loop's memory, so
// that no hipMalloc or hipFree (which in ROCm waits for the whole device) runs while solves run.
// Pieces are 256-byte aligned; a free piece merges with its free neighbours. Thread-safe.
class DeviceArena {
   public:
    explicit DeviceArena(size_t bytes) : size_((bytes + 255) / 256 * 256) {
        if (size_) FEM_HIP(hipMalloc((void**)&base_, size_));
        if (size_) free_[0] = size_;
    }
    ~DeviceArena() {
        if (base_) (void)hipFree(base_);
    }
    DeviceArena(const DeviceArena&) = delete;
    DeviceArena& operator=(const DeviceArena&) = delete;

    // The smallest free piece that holds `bytes` (best fit), or nullptr when none does.
    void* take(size_t bytes) {
        const size_t need = (bytes + 255) / 256 * 256;
        std::lock_guard<std::mutex> g(m_);
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= need && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) {
            ++misses_;
            return nullptr;
        }
        const size_t off = best->first, len = best->second;
        free_.erase(best);
        if (len > need) free_[off + need] = len - need;
        used_ += need;
        if (used_ > peak_) peak_ = used_;
        return base_ + off;
    }
    void give(void* p, size_t bytes) {
        const size_t need = (bytes + 255) / 256 * 256;
        std::lock_guard<std::mutex> g(m_);
        size_t off = (size_t)(static_cast<char*>(p) - base_), len = need;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[off] = len;
        used_ -= need;
    }
    size_t size() const { return size_; }
    size_t peak() {
        std::lock_guard<std::mutex> g(m_);
        return peak_;
    }
    // Requests that found no piece and went to hipMalloc instead.
    uint64_t misses() {
        std::lock_guard<std::mutex> g(m_);
        return misses_;
    }

   private:
    std::mutex m_;
    char* base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    uint64_t misses_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length of each free piece
};
// While set on a thread, the Buffers allocated there come from this arena (and go back to it when
// freed, from any thread); a request the arena cannot hold falls back to hipMalloc.
inline thread_local DeviceArena* alloc_arena = nullptr;

// The HIP objects of one solve: its stream, the pinned record its scalars are read back into, and
// its events. A serving loop keeps them in a pool (HandlePool) and hands them from one solve to
// the next, since creating and destroying them can wait for the whole device.
struct SolveHandles {
    hipStream_t s = nullptr;
    double* pinned = nullptr;  // 16 doubles
    hipEvent_t ev = nullptr, ev_sync = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;

    static SolveHandles create() {
        SolveHandles h;
        FEM_HIP(hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking));
        FEM_HIP(hipHostMalloc((void**)&h.pinned, 16 * sizeof(double), hipHostMallocDefault));
        FEM_HIP(hipEventCreateWithFlags(&h.ev, hipEventDisableTiming));
        FEM_HIP(hipEventCreateWithFlags(&h.ev_sync, hipEventDisableTiming));
        FEM_HIP(hipEventCreate(&h.ev_t0));
        FEM_HIP(hipEventCreate(&h.ev_t1));
        return h;
    }
    void destroy() {
        if (s) (void)hipStreamSynchronize(s);
        if (pinned) (void)hipHostFree(pinned);
        for (hipEvent_t e : {ev, ev_sync, ev_t0, ev_t1})
            if (e) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
        *this = SolveHandles{};
    }
};

// Handles given back by finished solves, taken by new ones (created when none is free). Thread-safe.
class HandlePool {
   public:
    HandlePool() = default;
    HandlePool(const HandlePool&) = delete;
    HandlePool& operator=(const HandlePool&) = delete;
    ~HandlePool() {
        for (SolveHandles& h : free_) h.destroy();
    }
    SolveHandles take() {
        {
            std::lock_guard<std::mutex> g(m_);
            if (!free_.empty()) {
                SolveHandles h = free_.back();
                free_.pop_back();
                return h;
            }
        }
        return SolveHandles::create();
    }
    // h's stream must be idle.
    void give(const SolveHandles& h) {
        std::lock_guard<std::mutex> g(m_);
        free_.push_back(h);
    }

   private:
    std::mutex m_;
    std::vector<SolveHandles> free_;
};
// While set on a thread, the solvers built there take their handles from this pool (and give them
// back when destroyed, on any thread).
inline thread_local HandlePool* solve_handles = nullptr;

class Buffer {
   public:
    Buffer() = default;
    explicit Buffer(size_t bytes) { alloc(bytes); }
    ~Buffer() { release(); }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    Buffer(Buffer&& o) noexcept : p_(o.p_), n_(o.n_), arena_(o.arena_), owned_(o.owned_) {
        o.p_ = nullptr;
        o.n_ = 0;
        o.arena_ = nullptr;
        o.owned_ = true;
    }
    Buffer& operator=(Buffer&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            arena_ = o.arena_;
            owned_ = o.owned_;
            o.p_ = nullptr;
            o.n_ = 0;
            o.arena_ = nullptr;
            o.owned_ = true;
        }
        return *this;
    }
    // Refers to another buffer's memory without owning it (the owner must outlive this view).
    void view(const Buffer& o) {
        release();
        p_ = o.p_;
        n_ = o.n_;
        owned_ = false;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes && alloc_arena) {
            p_ = alloc_arena->take(bytes);
            if (p_) arena_ = alloc_arena;
        }
        if (bytes && !p_) FEM_HIP(hipMalloc(&p_, bytes));
        n_ = bytes;
        if (alloc_meter) {
            alloc_meter->live += (int64_t)bytes;
            if (alloc_meter->live > alloc_meter->peak) alloc_meter->peak = alloc_meter->live;
        }
    }
    void release() {
        if (!owned_) {  // a view: nothing to free
            p_ = nullptr;
            n_ = 0;
            owned_ = true;
            return;
        }
        if (p_) {
            if (arena_) arena_->give(p_, n_);
            else (void)hipFree(p_);
        }
        if (alloc_meter) alloc_meter->live -= (int64_t)n_;
        p_ = nullptr;
        n_ = 0;
        arena_ = nullptr;
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
    DeviceArena* arena_ = nullptr;  // the arena the piece came from, or none (hipMalloc)
    bool owned_ = true;             // false for a view of another buffer's memory
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
    // This array refers to o's memory (o must outlive it).
    void view(const DArray& o) {
        buf_.view(o.buf_);
        n_ = o.n_;
    }
    // A zero-filled array of o's size.
    void like(const DArray& o, hipStream_t s) {
        resize(o.n_);
        if (n_) FEM_HIP(hipMemsetAsync(data(), 0, bytes(), s));
    }
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
