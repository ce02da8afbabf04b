/*
 * Test hooks of femahoy_amd: the building blocks of the dense coarse level, exposed so that the
 * test suite can check them against numpy. Not part of the solver boundary (fem_solver.h); no
 * counterpart in the reference.
 */
#ifndef FEM_TESTING_H
#define FEM_TESTING_H

#include <stddef.h>
#include <stdint.h>

#include "fem_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C = A B^T on the device's FP64 matrix cores (host arrays, row-major; A m x k, B n x k; m, n
 * multiples of 64, k a multiple of 16). */
int fem_test_gemm_nt(int m, int n, int k, const double* A, const double* B, double* C, char* err, size_t errlen);

/* The inverse of the symmetric positive definite n x n matrix A (host arrays) by the device's
 * blocked LDL^T and triangular inverse, in FP64. */
int fem_test_dense_inverse(int n, const double* A, double* Ainv, char* err, size_t errlen);

/* x = A^-1 b by the coarse solve of an FP32 V-cycle: the FP64 factor and inverse, the inverse
 * kept in FP32, the product summed in FP32. */
int fem_test_dense_solve32(int n, const double* A, const float* b, float* x, char* err, size_t errlen);

/* A plain copy of `bytes` (a multiple of 16) between two device buffers on the null stream, then
 * a wait: the calibration of HBM byte counters (FETCH_SIZE / WRITE_SIZE) against a kernel whose
 * traffic is known. */
int fem_test_copy(const void* d_src, void* d_dst, size_t bytes, char* err, size_t errlen);

/* The host setup of the grid multigrid of the assembled boundary (fem_solver_new_csr), alone: the
 * reference's level loop (multigrid.rs:84-139) with coarsen (329-452) and Galerkin products
 * (sparse.rs:233-292), or with `cells` (may be NULL) the coarse matrices of by_cells
 * (multigrid.rs:458-671), down to `coarsest` unknowns (max_levels > 0: at most that many levels).
 * No device is touched. */
typedef struct fem_csr_levels fem_csr_levels;
int fem_test_csr_levels_new(const fem_csr* a, const fem_grid_level* g, const fem_cells* cells, uint64_t coarsest,
                            int max_levels, fem_csr_levels** out, char* err, size_t errlen);
int fem_test_csr_levels_count(const fem_csr_levels* h);
/* Level l: its unknowns n, the entries of its matrix, and the columns and entries of the
 * prolongation from level l + 1 (0 and 0 on the last level). */
int fem_test_csr_level_sizes(const fem_csr_levels* h, int l, uint64_t* n, uint64_t* a_nnz, uint64_t* p_cols,
                             uint64_t* p_nnz);
/* Copies level l: the matrix (n + 1, a_nnz, a_nnz), the prolongation from level l + 1 (n + 1,
 * p_nnz, p_nnz; skipped on the last level) and the grid point (3 per unknown) and component of
 * each unknown. */
int fem_test_csr_level_arrays(const fem_csr_levels* h, int l, uint64_t* a_ptr, uint32_t* a_idx, double* a_val,
                              uint64_t* p_ptr, uint32_t* p_idx, double* p_val, uint32_t* points, uint8_t* comp);
void fem_test_csr_levels_free(fem_csr_levels* h);

/* The nested dissection order of a square pattern from the coordinates of its unknowns (3 each;
 * order.rs:19-53; order_out holds rows values, item k the old index of new unknown k), and the
 * counts of its symbolic factor (symbolic.rs:41-192: entries(), flops(), largest_front() and the
 * number of supernodes), by the library's host restatement that the FLOPS crossover uses. No
 * device is touched. */
int fem_test_symbolic(const fem_csr* a, const double* xyz, uint64_t* order_out, uint64_t* entries, double* flops,
                      uint64_t* largest_front, uint64_t* supernodes, char* err, size_t errlen);

/* The host part of fem_assemble alone: the reduced pattern (ptr: free + 1 values, idx: entries)
 * and the free unknowns of the mesh, with their counts first (call with NULL arrays to size
 * them). No device is touched. */
int fem_test_assembly_pattern(const fem_mesh* mesh, uint64_t* free_count, uint64_t* entries, uint64_t* ptr,
                              uint32_t* idx, int64_t* free_list, char* err, size_t errlen);

/* The structure of the multifrontal factor of a square pattern (symbolic.rs:41-150 in the nested
 * dissection order of the points xyz, 3 per unknown): the supernode and row counts first (call
 * with NULL arrays to size them), then the order (perm: n, new -> old), the supernodes' first
 * columns (start: supernodes + 1), their rows (rptr: supernodes + 1, rows: the factor's order,
 * each supernode's own columns first) and their parents (-1 at a root). join > 1: the supernodes
 * after the multifrontal factor's joins of at most `join` pivots (0: the fundamental ones). No
 * device is touched. */
int fem_test_supernodal(const fem_csr* a, const double* xyz, uint64_t join, uint64_t* nsup, uint64_t* nrows,
                        uint64_t* perm, uint64_t* start, uint64_t* rptr, uint64_t* rows, int64_t* sparent, char* err,
                        size_t errlen);

/* The condensed incompatible-mode hex of a box cell with sides size[3] and an isotropic material
 * (hex.rs:199-279 Hex::new's k: 24 x 24 row-major, unknown 3 corner + component, corners in VTK
 * order), by the library's host restatement that FEM_ELASTIC_HEX8 boxes assemble. No device. */
int fem_test_hex_element(const double* size, double e, double nu, double* out, char* err, size_t errlen);

/* The reduced system that fem_solver_new_box assembles for a FEM_ELASTIC_HEX8 box, on the host:
 * its counts first (call with NULL arrays to size them), then the reduced matrix (ptr: n + 1,
 * idx and val: nnz), the right side (n) and the largest aspect ratio of a cell. No device. */
int fem_test_hex_box_system(const fem_box* box, uint64_t* n, uint64_t* nnz, uint64_t* ptr, uint32_t* idx, double* val,
                            double* rhs, double* aspect, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* FEM_TESTING_H */
