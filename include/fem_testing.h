/*
 * Test hooks of femahoy_amd: the building blocks of the dense coarse level, exposed so that the
 * test suite can check them against numpy. Not part of the solver boundary (fem_solver.h); no
 * counterpart in the reference.
 */
#ifndef FEM_TESTING_H
#define FEM_TESTING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C = A B^T on the device's FP64 matrix cores (host arrays, row-major; A m x k, B n x k; m, n
 * multiples of 64, k a multiple of 16). */
int fem_test_gemm_nt(int m, int n, int k, const double* A, const double* B, double* C, char* err, size_t errlen);

/* The inverse of the symmetric positive definite n x n matrix A (host arrays) by the device's
 * blocked LDL^T and triangular inverse, in FP64. */
int fem_test_dense_inverse(int n, const double* A, double* Ainv, char* err, size_t errlen);

/* A plain copy of `bytes` (a multiple of 16) between two device buffers on the null stream, then
 * a wait: the calibration of HBM byte counters (FETCH_SIZE / WRITE_SIZE) against a kernel whose
 * traffic is known. */
int fem_test_copy(const void* d_src, void* d_dst, size_t bytes, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* FEM_TESTING_H */
