// A plain device copy with known traffic (read `bytes`, write `bytes`), the yardstick of the HBM
// byte counters.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/fem_testing.h"

namespace {

__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int fem_test_copy(const void* d_src, void* d_dst, size_t bytes, char* err, size_t errlen) {
    if (bytes % 16 != 0 || d_src == nullptr || d_dst == nullptr) {
        if (err && errlen) snprintf(err, errlen, "the copy needs two buffers and a multiple of 16 bytes");
        return 1;
    }
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)(n / 256 < 8192 ? (n + 255) / 256 : 8192);
    if (n > 0) hipLaunchKernelGGL(k_copy16, dim3(blocks ? blocks : 1), dim3(256), 0, 0, (const uint4*)d_src, (uint4*)d_dst, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (err && errlen) snprintf(err, errlen, "HIP error: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}
