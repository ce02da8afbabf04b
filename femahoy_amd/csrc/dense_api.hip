// The dense direct solve on device memory (include/fem_solver.h): the dense FP64 LDL^T of the
// coarse level and of the direct method (dense.hip), for a caller that holds a dense symmetric
// positive definite matrix on the device; and its allocation-free variant, whose workspace is a
// file-scope device array of fixed size.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/fem_solver.h"
#include "common.hpp"
#include "dense.hpp"

namespace {

void put(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}

template <class F>
int guard(char* err, size_t errlen, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        put(err, errlen, e.what());
        return 1;
    }
}

void check_args(int n, const double* a, const double* b, const double* x) {
    if (n < 0) throw fem::Error("the matrix needs a size of 0 or more");
    if (n > 0 && (!a || !b || !x)) throw fem::Error("the dense solve needs the matrix, the right side and x");
}

// The allocation-free variant: matrices of up to NOALLOC_MAX unknowns, whose workspace (the
// factor's, then one n-vector for the refinement residual) is `scratch`, 25.1 MiB that every
// device the library runs on holds from the load of the library's code until the process ends.
constexpr int NOALLOC_MAX = 1024;
constexpr size_t SCRATCH_BYTES =
    fem::DenseWork::bytes(NOALLOC_MAX) + fem::DenseWork::align((size_t)NOALLOC_MAX * sizeof(double));

}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

// The device address of `scratch` on each device, and a lock per device: calls on one device
// share its array, so they run one at a time.
struct ScratchSlot {
    void* ptr = nullptr;
    std::mutex m;
};
std::mutex slots_m;
std::map<int, ScratchSlot> slots;

// The slot of the current device; its address is taken on the first call on that device, with
// that device current (the host-side name `scratch` is not a device address).
ScratchSlot& scratch_slot() {
    int dev = 0;
    FEM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(slots_m);
    ScratchSlot& s = slots[dev];
    if (!s.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw fem::Error(std::string("the device address of the dense solve's workspace is not available: ") +
                             hipGetErrorString(e));
        s.ptr = p;
    }
    return s;
}

}  // namespace

extern "C" int fem_dense_solve_device_noalloc(int n, const double* d_a, const double* d_b, double* d_x, char* err,
                                              size_t errlen) {
    return guard(err, errlen, [&] {
        check_args(n, d_a, d_b, d_x);
        if (n > NOALLOC_MAX)
            throw fem::Error("the allocation-free dense solve holds at most " + std::to_string(NOALLOC_MAX) +
                             " unknowns; this matrix has " + std::to_string(n));
        if (n == 0) return;
        ScratchSlot& slot = scratch_slot();
        std::lock_guard<std::mutex> g(slot.m);
        const size_t wb = fem::DenseWork::bytes(n);  // at most bytes(NOALLOC_MAX): the size grows with n
        fem::DenseWork w;
        w.carve(slot.ptr, n);
        double* r = reinterpret_cast<double*>(static_cast<char*>(slot.ptr) + wb);
        hipStream_t s = nullptr;
        // every part of the workspace that is read is written first in this call (the padded
        // copy, the identity, the status cleared by a kernel), so the last call's contents do not
        // matter
        fem::dense_factor_inverse(d_a, n, w, s);
        fem::dense_solve_refined(d_a, n, w, d_b, d_x, r, s);
        FEM_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int fem_dense_solve_device(int n, const double* d_a, const double* d_b, double* d_x, char* err,
                                      size_t errlen) {
    return guard(err, errlen, [&] {
        check_args(n, d_a, d_b, d_x);
        if (n == 0) return;
        const size_t wb = fem::DenseWork::bytes(n);
        fem::Buffer work(wb + fem::DenseWork::align((size_t)n * sizeof(double)));
        fem::DenseWork w;
        w.carve(work.get(), n);
        double* r = reinterpret_cast<double*>(static_cast<char*>(work.get()) + wb);
        hipStream_t s = nullptr;
        fem::dense_factor_inverse(d_a, n, w, s);
        fem::dense_solve_refined(d_a, n, w, d_b, d_x, r, s);
        FEM_HIP(hipStreamSynchronize(s));  // the workspace is freed at the end of this scope
    });
}
