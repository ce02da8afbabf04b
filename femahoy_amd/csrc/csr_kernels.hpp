// Device kernels of the assembled (CSR) path: the operator of each level, and the transfers (the
// prolongation and its transpose), are sparse matrices (FP64 for the outer PCG, the V-cycle's
// precision below it).
//
// On the device a matrix is kept in sliced ELL form: rows in slices of 64 (one wavefront), each
// slice padded to its longest row and stored column-major (entry j of the slice's 64 rows is 64
// consecutive values), so that every load of a wave reads one contiguous run. A thread sums its
// row's entries in their CSR order; padding entries have value 0 and point at the slice's smallest
// column. Offsets are 32-bit: the boundary refuses more than 2^32 - 1 padded entries per matrix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fem {

constexpr int SELL_C = 64;  // rows per slice

// A matrix in sliced ELL form. Columns are C: uint32_t, or (when every slice's columns span less
// than 2^16) uint16_t offsets from the slice's smallest column, cfirst[slice], which cuts the bytes
// of an FP32 entry from 8 to 6.
template <class V, class C>
struct SellMat {
    using col_t = C;
    const uint32_t* base;    // slice s: entries base[s] .. base[s + 1]
    const uint32_t* cfirst;  // uint16_t columns: the slice's smallest column
    const C* col;
    const V* val;
};

// The sum of row `row` (slice base `b`, `width` entries per row, columns from c0) times x, in the
// row's order. The entries go in groups of SELL_U: all their columns and values are loaded, then all
// their x, before the group is summed, so that a long row (a restriction's, or any row of a small
// level, where no other wave hides the latency) waits for memory once per group, not once per entry.
constexpr uint32_t SELL_U = 8;
template <class T, class V, class C>
__device__ __forceinline__ T sell_row(const SellMat<V, C>& m, const T* __restrict__ x, uint32_t b, uint32_t width,
                                      int lane, uint32_t c0) {
    const C* __restrict__ col = m.col;
    const V* __restrict__ val = m.val;
    T s = T(0);
    uint32_t k = b + lane;
    uint32_t j = 0;
    for (; j + SELL_U <= width; j += SELL_U, k += SELL_U * SELL_C) {
        uint32_t c[SELL_U];
        V v[SELL_U];
        T xv[SELL_U];
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) {
            c[u] = c0 + (uint32_t)col[k + u * SELL_C];
            v[u] = val[k + u * SELL_C];
        }
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) s += (T)v[u] * xv[u];
    }
    for (; j < width; ++j, k += SELL_C) s += (T)val[k] * x[c0 + (uint32_t)col[k]];
    return s;
}

// This thread's row, its lane in the slice, and the slice's base, width and first column.
template <class V, class C>
__device__ __forceinline__ bool sell_slot(uint64_t rows, const SellMat<V, C>& m, uint32_t& row, int& lane, uint32_t& b,
                                          uint32_t& w, uint32_t& c0) {
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    row = (uint32_t)t;
    lane = (int)(t % SELL_C);
    const uint64_t slice = t / SELL_C;
    if (slice * SELL_C >= rows) return false;  // the whole slice lies past the last row
    b = m.base[slice];
    w = (m.base[slice + 1] - b) / SELL_C;
    c0 = sizeof(C) == 2 ? m.cfirst[slice] : 0u;
    return true;
}

// y = A x.
template <class T, class V, class C>
__global__ void __launch_bounds__(BLOCK) k_csr_apply(uint64_t n, SellMat<V, C> m, const T* __restrict__ x,
                                                     T* __restrict__ y) {
    uint32_t row, b, w, c0;
    int lane;
    if (!sell_slot(n, m, row, lane, b, w, c0)) return;
    const T s = sell_row<T>(m, x, b, w, lane, c0);
    if (row < n) y[row] = s;
}

// One Chebyshev step of the smoother, as k_cheb: d = c1 d + c2 D^-1 (b - A x), x2 = x + d.
template <class T, class V, class C, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_cheb(uint64_t n, SellMat<V, C> m, const T* __restrict__ x,
                                                    T* __restrict__ x2, T* __restrict__ d, const T* __restrict__ b,
                                                    const S* __restrict__ invd, T c1, T c2) {
    uint32_t row, sb, w, c0;
    int lane;
    if (!sell_slot(n, m, row, lane, sb, w, c0)) return;
    const T s = sell_row<T>(m, x, sb, w, lane, c0);
    if (row < n) {
        const T r = b[row] - s;
        const T dn = c1 * d[row] + c2 * ld<T>(invd, row) * r;
        d[row] = dn;
        x2[row] = x[row] + dn;
    }
}

// r = b - A x.
template <class T, class V, class C>
__global__ void __launch_bounds__(BLOCK) k_csr_residual(uint64_t n, SellMat<V, C> m, const T* __restrict__ x,
                                                        const T* __restrict__ b, T* __restrict__ r) {
    uint32_t row, sb, w, c0;
    int lane;
    if (!sell_slot(n, m, row, lane, sb, w, c0)) return;
    const T s = sell_row<T>(m, x, sb, w, lane, c0);
    if (row < n) r[row] = b[row] - s;
}

// b_c = R r_f with R = P^T; with invd, also the coarse level's first Chebyshev step from x = 0,
// d_c = x_c = c0 D^-1 b_c (as k_cheb_first).
template <class T, class V, class C, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_restrict(uint64_t nc, SellMat<V, C> m, const T* __restrict__ rf,
                                                        T* __restrict__ bc, const S* __restrict__ invd, T c0,
                                                        T* __restrict__ xc, T* __restrict__ dc) {
    uint32_t row, sb, w, cf;
    int lane;
    if (!sell_slot(nc, m, row, lane, sb, w, cf)) return;
    const T s = sell_row<T>(m, rf, sb, w, lane, cf);
    if (row < nc) {
        bc[row] = s;
        if (invd) {
            const T v = c0 * ld<T>(invd, row) * s;
            dc[row] = v;
            xc[row] = v;
        }
    }
}

// x_f += P x_c.
template <class T, class V, class C>
__global__ void __launch_bounds__(BLOCK) k_csr_prolong_add(uint64_t nf, SellMat<V, C> m, const T* __restrict__ xc,
                                                           T* __restrict__ xf) {
    uint32_t row, sb, w, c0;
    int lane;
    if (!sell_slot(nf, m, row, lane, sb, w, c0)) return;
    const T s = sell_row<T>(m, xc, sb, w, lane, c0);
    if (row < nf) xf[row] += s;
}

// q = A p (FP64) and the block partials of p.q.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_q(uint64_t n, SellMat<double, C> m, const double* __restrict__ p,
                                                     double* __restrict__ q, double* __restrict__ part) {
    uint32_t row, sb, w, c0;
    int lane;
    double pq = 0.0;
    if (sell_slot(n, m, row, lane, sb, w, c0)) {
        const double s = sell_row<double>(m, p, sb, w, lane, c0);
        if (row < n) {
            q[row] = s;
            pq = p[row] * s;
        }
    }
    pq = block_sum(pq);
    if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

// r = b - A x (FP64) and the block partials of b.b and r.r.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_start(uint64_t n, SellMat<double, C> m,
                                                         const double* __restrict__ x, const double* __restrict__ b,
                                                         double* __restrict__ r, double* __restrict__ part_bb,
                                                         double* __restrict__ part_rr) {
    uint32_t row, sb, w, c0;
    int lane;
    double bb = 0.0, rr = 0.0;
    if (sell_slot(n, m, row, lane, sb, w, c0)) {
        const double s = sell_row<double>(m, x, sb, w, lane, c0);
        if (row < n) {
            const double rv = b[row] - s;
            r[row] = rv;
            bb = b[row] * b[row];
            rr = rv * rv;
        }
    }
    bb = block_sum(bb);
    rr = block_sum(rr);
    if (threadIdx.x == 0) {
        part_bb[blockIdx.x] = bb;
        part_rr[blockIdx.x] = rr;
    }
}

// The rows of a CSR (n rows, columns < n) into a dense row-major n x n matrix (zero elsewhere).
__global__ void __launch_bounds__(BLOCK) k_csr_scatter_dense(uint64_t n, const uint64_t* __restrict__ ptr,
                                                             const uint32_t* __restrict__ idx,
                                                             const double* __restrict__ val, double* __restrict__ A) {
    const uint64_t r = blockIdx.x * (uint64_t)BLOCK + threadIdx.x;
    if (r >= n) return;
    for (uint64_t k = ptr[r]; k < ptr[r + 1]; ++k) A[r * n + idx[k]] = val[k];
}

}  // namespace fem
