// The dense coarse level: an FP64 LDL^T factor on the FP64 matrix cores (MFMA), the explicit
// inverse A^-1 = L^-T D^-1 L^-1 formed from it, and the per-cycle solve as a matrix-vector
// product with that inverse, in the V-cycle's precision.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "common.hpp"

namespace fem {

// The most unknowns of a dense direct solve (the direct method of both boundaries, and the linear
// level of a p-multigrid with no grid below it): the reference's TRANSIENT_DIRECT_LIMIT
// (crates/fem/src/solve.rs:24), its largest direct crossover. A factor of n unknowns holds four
// n^2 FP64 arrays on the device, 51 GB at this limit.
constexpr uint64_t DENSE_DIRECT_LIMIT = 40000;
// The direct method of the assembled boundary factors larger matrices by the multifrontal LDL^T.
constexpr uint64_t MULTIFRONTAL_FROM = 8000;

// The device workspace of one dense factor of n unknowns (np = n rounded up to 64, nb = np / 64
// blocks), carved from one contiguous region of bytes(n) bytes at 256-byte aligned offsets:
// A (np x np, the padded copy that the factor overwrites with L and D), U (np x np, L^-T),
// inv (np x np, the FP64 inverse), P (np x 64, the scaled block column of a step), linv
// (nb blocks of 64 x 64, the inverse of each diagonal block's L), bk (64 x 64), d and dg (np:
// the pivots and the original diagonal) and status (the first failed pivot's column + 1, or 0).
struct DenseWork {
    int n = 0, np = 0, nb = 0;
    double *A = nullptr, *U = nullptr, *inv = nullptr, *P = nullptr, *linv = nullptr, *bk = nullptr, *d = nullptr,
           *dg = nullptr;
    int* status = nullptr;

    static constexpr size_t align(size_t b) { return (b + 255) / 256 * 256; }
    static constexpr size_t bytes(int n) {
        const size_t np = (size_t)(n + 63) / 64 * 64, nb = np / 64;
        return 3 * align(np * np * 8) + align(np * 64 * 8) + align(nb * 64 * 64 * 8) + align(64 * 64 * 8) +
               2 * align(np * 8) + align(sizeof(int));
    }
    // Sets the pointers for n unknowns from `base` (bytes(n) bytes of device memory, 256-byte aligned).
    void carve(void* base, int n_);
};

// Factors the n x n symmetric positive definite matrix `a` (FP64, row-major, leading dimension n,
// device memory, not changed) into the workspace and forms the FP64 inverse in w.inv (np x np,
// the padding rows and columns those of the identity). Waits for the stream (the pivot test) and
// throws fem::Error with the reference's "not positive definite" text when a pivot fails.
// Allocates nothing.
void dense_factor_inverse(const double* a, int n, const DenseWork& w, hipStream_t s);

// One step of the blocked right-looking LDL^T of the row-major matrix A (lda, nb blocks of 64 per
// side, lower triangle): factor diagonal block kb (its L into the block, D into d[64 kb ..], its
// L^-1 into linv[kb]; a pivot at or below 1e-10 |dg| sets *status to its column + 1 when *status
// is 0), form the block column below it, and subtract its product from every block (i, j) with
// kb < j <= i. bk (64 x 64) and P (64 (nb - kb - 1) x 64) are workspace.
void ldlt_panel(double* A, int lda, int nb, int kb, const double* dg, double* d, double* linv, double* bk, double* P,
                int* status, hipStream_t s);

// x = A^-1 b with one step of refinement on the FP64 residual of `a`: x = inv b, r = b - A x,
// x += inv r. inv is w.inv (FP64, leading dimension w.np); r is an n-vector workspace. Allocates
// nothing.
void dense_solve_refined(const double* a, int n, const DenseWork& w, const double* b, double* x, double* r,
                         hipStream_t s);

class DenseCoarse {
   public:
    // Factors the n x n symmetric matrix `a` (FP64, row-major, device memory, overwritten) and
    // forms its inverse. `fp32` keeps the inverse in FP32 for an FP32 V-cycle. Throws fem::Error
    // with the reference's "not positive definite" text when a pivot fails.
    void factor(double* a, int n, bool fp32, hipStream_t s);

    // The factor of o for another solve of the same level: its inverse and maps viewed (o must
    // outlive this one), its dense right side of its own (zero).
    void shadow_of(const DenseCoarse& o, hipStream_t s);

    // x[map[i]] = sum_j inv[i][j] b[map[j]] for the n dense unknowns; map holds the grid unknown
    // of each dense unknown. T is float or double (the storage of the inverse must match).
    template <class T>
    void solve(const T* b, T* x, hipStream_t s) const;

    // The dense right side (n values, zero padded) that solve() gathers b into: a producer may
    // write it directly (rhs()[dense index] = value) and call solve_rhs(), skipping the gather.
    template <class T>
    T* rhs() const;
    template <class T>
    void solve_rhs(T* x, hipStream_t s) const;
    // The dense index of each unknown of the level, or -1 (the inverse of map), for producers of
    // rhs(); empty unless set by the owner.
    DArray<int32_t>& dense_of() { return dense_of_; }
    const DArray<int32_t>& dense_of() const { return dense_of_; }

    int n() const { return n_; }
    // The device arrays of the per-cycle solve (the inverse and the dense right side).
    const void* inverse_data() const { return fp32_ ? (const void*)inv32_.data() : (const void*)inv64_.data(); }
    const void* rhs_data() const { return fp32_ ? (const void*)bc32_.data() : (const void*)bc64_.data(); }
    // Algorithmic bytes of one solve in storage of vb bytes: the inverse's rows, the dense right
    // side and the map in, x out; with `gather` also the gather (map, b in, the dense right side out).
    uint64_t solve_bytes(size_t vb, bool gather) const {
        const uint64_t n = (uint64_t)n_, ld = (uint64_t)ld_;
        return n * ld * vb + ld * vb + n * 8 + n * vb + (gather ? n * 8 + 2 * n * vb : 0);
    }
    const DArray<int64_t>& map() const { return map_; }
    DArray<int64_t>& map() { return map_; }
    double factor_seconds() const { return factor_seconds_; }
    double factor_flops() const { return factor_flops_; }

   private:
    int n_ = 0, np_ = 0, ld_ = 0;
    bool fp32_ = false;
    DArray<double> inv64_, bc64_;
    DArray<float> inv32_, bc32_;
    DArray<int64_t> map_;
    DArray<int32_t> dense_of_;
    double factor_seconds_ = 0.0, factor_flops_ = 0.0;
};

// The FP64 matrix-core tile product used by the factor, exposed for its unit test:
// C = alpha C + beta A B^T with A (m x k), B (n x k), C (m x n), all row-major, m, n, k
// multiples of 64.
void dense_gemm_nt(int m, int n, int k, double alpha, const double* A, int lda, double beta, const double* B, int ldb,
                   double* C, int ldc, hipStream_t s);

}  // namespace fem
