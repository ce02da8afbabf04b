// The dense direct solve on device memory (include/fem_solver.h): the dense FP64 LDL^T of the
// coarse level and of the direct method (dense.hip), for a caller that holds a dense symmetric
// positive definite matrix on the device.
#include <hip/hip_runtime.h>

#include <cstring>
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

}  // namespace

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
