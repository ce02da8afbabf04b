// Device kernels of the assembled (CSR) path: the operator of each level is a CSR matrix (FP64
// for the outer PCG, the V-cycle's precision below it), the transfers are the prolongation and
// its transpose as CSR matrices.
//
// A row is summed by a group of L lanes of a wavefront (L a power of 2 dividing 64, chosen from
// the mean row length): lane j takes entries j, j + L, ... of the row, so that a group reads its
// row's columns and values in contiguous runs, and the L partial sums are combined by butterfly
// shuffles in a fixed order (the result does not depend on timing). The row index and the offsets
// are 32-bit: the boundary refuses more than 2^32 - 1 entries per matrix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fem {

// The sum of row `row` of (ptr, idx, val) times x, by the L lanes of a group (lane = lane in the
// group); valid in every lane of the group.
template <int L, class T, class V>
__device__ __forceinline__ T csr_row(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ idx,
                                     const V* __restrict__ val, const T* __restrict__ x, uint32_t row, bool live,
                                     int lane) {
    T s = T(0);
    if (live) {
        const uint32_t e = ptr[row + 1];
        for (uint32_t k = ptr[row] + lane; k < e; k += L) s += (T)val[k] * x[idx[k]];
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, L);
    return s;
}

// The row and lane of this thread for groups of L lanes.
template <int L>
__device__ __forceinline__ void csr_slot(uint64_t n, uint32_t& row, int& lane, bool& live) {
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    row = (uint32_t)(t / L);
    lane = (int)(t % L);
    live = t / L < n;
}

// y = A x.
template <int L, class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_apply(uint64_t n, const uint32_t* __restrict__ ptr,
                                                     const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                                     const T* __restrict__ x, T* __restrict__ y) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(n, row, lane, live);
    const T s = csr_row<L, T, V>(ptr, idx, val, x, row, live, lane);
    if (live && lane == 0) y[row] = s;
}

// One Chebyshev step of the smoother, as k_cheb: d = c1 d + c2 D^-1 (b - A x), x2 = x + d.
template <int L, class T, class V, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_cheb(uint64_t n, const uint32_t* __restrict__ ptr,
                                                    const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                                    const T* __restrict__ x, T* __restrict__ x2, T* __restrict__ d,
                                                    const T* __restrict__ b, const S* __restrict__ invd, T c1, T c2) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(n, row, lane, live);
    const T s = csr_row<L, T, V>(ptr, idx, val, x, row, live, lane);
    if (live && lane == 0) {
        const T r = b[row] - s;
        const T dn = c1 * d[row] + c2 * ld<T>(invd, row) * r;
        d[row] = dn;
        x2[row] = x[row] + dn;
    }
}

// r = b - A x.
template <int L, class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_residual(uint64_t n, const uint32_t* __restrict__ ptr,
                                                        const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                                        const T* __restrict__ x, const T* __restrict__ b,
                                                        T* __restrict__ r) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(n, row, lane, live);
    const T s = csr_row<L, T, V>(ptr, idx, val, x, row, live, lane);
    if (live && lane == 0) r[row] = b[row] - s;
}

// b_c = R r_f with R = P^T; with invd, also the coarse level's first Chebyshev step from x = 0,
// d_c = x_c = c0 D^-1 b_c (as k_cheb_first).
template <int L, class T, class V, class S>
__global__ void __launch_bounds__(BLOCK) k_csr_restrict(uint64_t nc, const uint32_t* __restrict__ ptr,
                                                        const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                                        const T* __restrict__ rf, T* __restrict__ bc,
                                                        const S* __restrict__ invd, T c0, T* __restrict__ xc,
                                                        T* __restrict__ dc) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(nc, row, lane, live);
    const T s = csr_row<L, T, V>(ptr, idx, val, rf, row, live, lane);
    if (live && lane == 0) {
        bc[row] = s;
        if (invd) {
            const T v = c0 * ld<T>(invd, row) * s;
            dc[row] = v;
            xc[row] = v;
        }
    }
}

// x_f += P x_c.
template <int L, class T, class V>
__global__ void __launch_bounds__(BLOCK) k_csr_prolong_add(uint64_t nf, const uint32_t* __restrict__ ptr,
                                                           const uint32_t* __restrict__ idx, const V* __restrict__ val,
                                                           const T* __restrict__ xc, T* __restrict__ xf) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(nf, row, lane, live);
    const T s = csr_row<L, T, V>(ptr, idx, val, xc, row, live, lane);
    if (live && lane == 0) xf[row] += s;
}

// q = A p (FP64) and the block partials of p.q.
template <int L>
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_q(uint64_t n, const uint32_t* __restrict__ ptr,
                                                     const uint32_t* __restrict__ idx, const double* __restrict__ val,
                                                     const double* __restrict__ p, double* __restrict__ q,
                                                     double* __restrict__ part) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(n, row, lane, live);
    const double s = csr_row<L, double, double>(ptr, idx, val, p, row, live, lane);
    double pq = 0.0;
    if (live && lane == 0) {
        q[row] = s;
        pq = p[row] * s;
    }
    pq = block_sum(pq);
    if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

// r = b - A x (FP64) and the block partials of b.b and r.r.
template <int L>
__global__ void __launch_bounds__(BLOCK) k_csr_pcg_start(uint64_t n, const uint32_t* __restrict__ ptr,
                                                         const uint32_t* __restrict__ idx,
                                                         const double* __restrict__ val, const double* __restrict__ x,
                                                         const double* __restrict__ b, double* __restrict__ r,
                                                         double* __restrict__ part_bb, double* __restrict__ part_rr) {
    uint32_t row;
    int lane;
    bool live;
    csr_slot<L>(n, row, lane, live);
    const double s = csr_row<L, double, double>(ptr, idx, val, x, row, live, lane);
    double bb = 0.0, rr = 0.0;
    if (live && lane == 0) {
        const double rv = b[row] - s;
        r[row] = rv;
        bb = b[row] * b[row];
        rr = rv * rv;
    }
    bb = block_sum(bb);
    rr = block_sum(rr);
    if (threadIdx.x == 0) {
        part_bb[blockIdx.x] = bb;
        part_rr[blockIdx.x] = rr;
    }
}

// v[i] = (T) a[i] for the values of a matrix (FP64 to the V-cycle's precision).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_csr_values(uint64_t n, const double* __restrict__ a, T* __restrict__ v) {
    const uint64_t u = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (u < n) v[u] = (T)a[u];
}

}  // namespace fem
