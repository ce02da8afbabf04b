// Device kernels of the assembled (CSR) path: the operator of each level, and the transfers (the
// prolongation and its transpose), are sparse matrices (FP64 for the outer PCG, the V-cycle's
// precision below it).
//
// On the device a matrix is kept in sliced ELL form: rows in slices of 64 (one wavefront), each
// slice padded to its longest row and stored column-major (entry j of the slice's 64 rows is 64
// consecutive values), so that every load of a wave reads one contiguous run. A thread sums its
// row's entries in their CSR order; padding entries have value 0 and point at the row itself.
// Offsets are 32-bit: the boundary refuses more than 2^32 - 1 padded entries per matrix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fem {

constexpr int SELL_C = 64;  // rows per slice

// The sum of row `row` (slice base `base`, `width` entries per row) times x, in the row's order.
// The entries go in groups of SELL_U: all their columns and values are loaded, then all their x,
// before the group is summed, so that a long row (a restriction's, or any row of a small level,
// where no other wave hides the latency) waits for memory once per group, not once per entry.
constexpr uint32_t SELL_U = 8;
template <class T, class V>
__device__ __forceinline__ T sell_row(const uint32_t* __restrict__ col, const V* __restrict__ val,
                                      const T* __restrict__ x, uint32_t base, uint32_t width, int lane) {
    T s = T(0);
    uint32_t k = base + lane;
    uint32_t j = 0;
    for (; j + SELL_U <= width; j += SELL_U, k += SELL_U * SELL_C) {
        uint32_t c[SELL_U];
        V v[SELL_U];
        T xv[SELL_U];
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) {
            c[u] = col[k + u * SELL_C];
            v[u] = val[k + u * SELL_C];
        }
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (uint32_t u = 0; u < SELL_U; ++u) s += (T)v[u] * xv[u];
    }
    for (; j < width; ++j, k += SELL_C) s += (T)val[k] * x[col[k]];
    return s;
}

// This thread's row, its lane in the slice, and the slice's base and width.
__device__ __forceinline__ bool sell_slot(uint64_t rows, const uint32_t* __restrict__ base, uint32_t& row, int& lane,
                                          uint32_t& b, uint32_t& w) {
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    row = (uint32_t)t;
    lane = (int)(t % SELL_C);
    const uint64_t slice = t / SELL_C;
    if (slice * SELL_C >= rows) return false;  // the whole slice lies past the last row
    b = base[slice];
    w = (base[slice + 1] - b) / SELL_C;
    return true;
}

// y = A x.
template <class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_apply(uint64_t n, const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ col, const V* __restrict__ val,
                                                     const T* __restrict__ x, T* __restrict__ y) {
    uint32_t row, b, w;
    int lane;
    if (!sell_slot(n, base, row, lane, b, w)) return;
    const T s = sell_row<T, V>(col, val, x, b, w, lane);
    if (row < n) y[row] = s;
}

// One Chebyshev step of the smoother, as k_cheb: d = c1 d + c2 D^-1 (b - A x), x2 = x + d.
template <class T, class V, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_cheb(uint64_t n, const uint32_t* __restrict__ base,
                                                    const uint32_t* __restrict__ col, const V* __restrict__ val,
                                                    const T* __restrict__ x, T* __restrict__ x2, T* __restrict__ d,
                                                    const T* __restrict__ b, const S* __restrict__ invd, T c1, T c2) {
    uint32_t row, sb, w;
    int lane;
    if (!sell_slot(n, base, row, lane, sb, w)) return;
    const T s = sell_row<T, V>(col, val, x, sb, w, lane);
    if (row < n) {
        const T r = b[row] - s;
        const T dn = c1 * d[row] + c2 * ld<T>(invd, row) * r;
        d[row] = dn;
        x2[row] = x[row] + dn;
    }
}

// r = b - A x.
template <class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_residual(uint64_t n, const uint32_t* __restrict__ base,
                                                        const uint32_t* __restrict__ col, const V* __restrict__ val,
                                                        const T* __restrict__ x, const T* __restrict__ b,
                                                        T* __restrict__ r) {
    uint32_t row, sb, w;
    int lane;
    if (!sell_slot(n, base, row, lane, sb, w)) return;
    const T s = sell_row<T, V>(col, val, x, sb, w, lane);
    if (row < n) r[row] = b[row] - s;
}

// b_c = R r_f with R = P^T; with invd, also the coarse level's first Chebyshev step from x = 0,
// d_c = x_c = c0 D^-1 b_c (as k_cheb_first).
template <class T, class V, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_restrict(uint64_t nc, const uint32_t* __restrict__ base,
                                                        const uint32_t* __restrict__ col, const V* __restrict__ val,
                                                        const T* __restrict__ rf, T* __restrict__ bc,
                                                        const S* __restrict__ invd, T c0, T* __restrict__ xc,
                                                        T* __restrict__ dc) {
    uint32_t row, sb, w;
    int lane;
    if (!sell_slot(nc, base, row, lane, sb, w)) return;
    const T s = sell_row<T, V>(col, val, rf, sb, w, lane);
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
template <class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_prolong_add(uint64_t nf, const uint32_t* __restrict__ base,
                                                           const uint32_t* __restrict__ col, const V* __restrict__ val,
                                                           const T* __restrict__ xc, T* __restrict__ xf) {
    uint32_t row, sb, w;
    int lane;
    if (!sell_slot(nf, base, row, lane, sb, w)) return;
    const T s = sell_row<T, V>(col, val, xc, sb, w, lane);
    if (row < nf) xf[row] += s;
}

// q = A p (FP64) and the block partials of p.q.
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_q(uint64_t n, const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ p, double* __restrict__ q,
                                                     double* __restrict__ part) {
    uint32_t row, sb, w;
    int lane;
    double pq = 0.0;
    if (sell_slot(n, base, row, lane, sb, w)) {
        const double s = sell_row<double, double>(col, val, p, sb, w, lane);
        if (row < n) {
            q[row] = s;
            pq = p[row] * s;
        }
    }
    pq = block_sum(pq);
    if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

// r = b - A x (FP64) and the block partials of b.b and r.r.
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_start(uint64_t n, const uint32_t* __restrict__ base,
                                                         const uint32_t* __restrict__ col,
                                                         const double* __restrict__ val, const double* __restrict__ x,
                                                         const double* __restrict__ b, double* __restrict__ r,
                                                         double* __restrict__ part_bb, double* __restrict__ part_rr) {
    uint32_t row, sb, w;
    int lane;
    double bb = 0.0, rr = 0.0;
    if (sell_slot(n, base, row, lane, sb, w)) {
        const double s = sell_row<double, double>(col, val, x, sb, w, lane);
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

}  // namespace fem
