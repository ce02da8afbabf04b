// Kernels of the solve: operator applies fused with the Chebyshev-Jacobi smoother, residuals,
// grid transfers, FP64 PCG vector updates with block-partial dot products, and conversions.
// Every reduction is two-stage with a fixed order (block partials, then one block sums them), so
// a result is bitwise the same from run to run and alone or next to other solves.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ops.hpp"

namespace fem {

constexpr int BLOCK = 256;

// Operators with a uniform interface: C components per node, apply and diagonal of a node row.
template <class Op>
struct ScalarOp {
    static constexpr int C = 1;
    Op op;
    __device__ __forceinline__ GridDims dims() const { return op.g; }
    template <class T>
    __device__ __forceinline__ void apply(const T* x, int i, int j, int k, T* out) const {
        out[0] = op.apply(x, i, j, k);
    }
    template <class T>
    __device__ __forceinline__ void diag(int i, int j, int k, T* out) const {
        out[0] = op.diag(i, j, k);
    }
};

template <class Op>
struct VectorOp {
    static constexpr int C = 3;
    Op op;
    __device__ __forceinline__ GridDims dims() const { return op.g; }
    template <class T>
    __device__ __forceinline__ void apply(const T* x, int i, int j, int k, T* out) const {
        op.apply3(x, i, j, k, out);
    }
    template <class T>
    __device__ __forceinline__ void diag(int i, int j, int k, T* out) const {
        op.diag3(i, j, k, out);
    }
};

__device__ __forceinline__ void node_ijk(const GridDims& g, size_t node, int& i, int& j, int& k) {
    i = (int)(node % (size_t)g.nx);
    const size_t r = node / (size_t)g.nx;
    j = (int)(r % (size_t)g.ny);
    k = (int)(r / (size_t)g.ny);
}

// Sum of v over a block of NT threads, deterministic; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum_t(double v) {
    __shared__ double part[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) s += part[w];
    __syncthreads();  // `part` may be written again by a following call
    return s;
}

__device__ __forceinline__ double block_sum(double v) { return block_sum_t<BLOCK>(v); }

// The block size of the kernels that sum block partials.
constexpr int RBLOCK = 1024;

// ------------------------------------------------------------------ operator kernels
// The logical block of a 1-D launch, XCD-aware: workgroups go to the 8 XCDs round-robin by id
// and each XCD has its own L2, so id p takes block (p % 8) * (G / 8) + p / 8 and neighbouring
// blocks, which read each other's halo, share an L2. The tail beyond a multiple of 8 keeps its id.
__device__ __forceinline__ size_t xcd_block(unsigned p, unsigned G) {
    const unsigned per = G / 8;
    return p < per * 8 ? (size_t)(p % 8) * per + p / 8 : (size_t)p;
}
__device__ __forceinline__ size_t xcd_block() { return xcd_block(blockIdx.x, gridDim.x); }

// y = A x on the unknowns; 0 on fixed nodes.
template <class OP, class T>
__global__ void __launch_bounds__(BLOCK) k_apply(OP op, const T* __restrict__ x, T* __restrict__ y) {
    const GridDims g = op.dims();
    const size_t node = xcd_block() * BLOCK + threadIdx.x;
    if (node >= g.nodes()) return;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    T out[OP::C];
    if (is_fixed(g, i, j, k)) {
#pragma unroll
        for (int c = 0; c < OP::C; ++c) out[c] = T(0);
    } else {
        op.apply(x, i, j, k, out);
    }
#pragma unroll
    for (int c = 0; c < OP::C; ++c) y[OP::C * node + c] = out[c];
}

// r = b - A x.
template <class OP, class T>
__device__ __forceinline__ void residual_body(unsigned blk, unsigned nblk, OP op, const T* __restrict__ x, const T* __restrict__ b,
                                                    T* __restrict__ r) {
    const GridDims g = op.dims();
    const size_t node = xcd_block(blk, nblk) * BLOCK + threadIdx.x;
    if (node >= g.nodes()) return;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    T out[OP::C];
    if (is_fixed(g, i, j, k)) {
#pragma unroll
        for (int c = 0; c < OP::C; ++c) r[OP::C * node + c] = T(0);
        return;
    }
    op.apply(x, i, j, k, out);
#pragma unroll
    for (int c = 0; c < OP::C; ++c) r[OP::C * node + c] = b[OP::C * node + c] - out[c];
}
template <class OP, class T>
__global__ void __launch_bounds__(BLOCK) k_residual(OP op, const T* __restrict__ x, const T* __restrict__ b,
                                                    T* __restrict__ r) {
    residual_body<OP, T>(blockIdx.x, gridDim.x, op, x, b, r);
}
template <class OP, class T>
struct KResidual {
    static constexpr int NT = BLOCK, MINB = 0;
    OP op;
    const T* x;
    const T* b;
    T* r;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        residual_body<OP, T>(blk, nblk, op, x, b, r);
    }
};

// One Chebyshev step from a nonzero x: r = b - A x_in, d = c1 d + c2 D^-1 r, x_out = x_in + d.
// D^-1 stored in S (T, or BF16).
template <class OP, class T, class S = T>
__device__ __forceinline__ void cheb_body(unsigned blk, unsigned nblk, OP op, const T* __restrict__ xin, T* __restrict__ xout,
                                                T* __restrict__ d, const T* __restrict__ b,
                                                const S* __restrict__ invd, T c1, T c2) {
    const GridDims g = op.dims();
    const size_t node = xcd_block(blk, nblk) * BLOCK + threadIdx.x;
    if (node >= g.nodes()) return;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    if (is_fixed(g, i, j, k)) {
#pragma unroll
        for (int c = 0; c < OP::C; ++c) {
            xout[OP::C * node + c] = T(0);
            d[OP::C * node + c] = T(0);
        }
        return;
    }
    T out[OP::C];
    op.apply(xin, i, j, k, out);
#pragma unroll
    for (int c = 0; c < OP::C; ++c) {
        const size_t u = OP::C * node + c;
        const T r = b[u] - out[c];
        const T dn = c1 * d[u] + c2 * ld<T>(invd, u) * r;
        d[u] = dn;
        xout[u] = xin[u] + dn;
    }
}
template <class OP, class T, class S = T>
__global__ void __launch_bounds__(BLOCK) k_cheb(OP op, const T* __restrict__ xin, T* __restrict__ xout,
                                                T* __restrict__ d, const T* __restrict__ b,
                                                const S* __restrict__ invd, T c1, T c2) {
    cheb_body<OP, T, S>(blockIdx.x, gridDim.x, op, xin, xout, d, b, invd, c1, c2);
}
template <class OP, class T, class S = T>
struct KCheb {
    static constexpr int NT = BLOCK, MINB = 0;
    OP op;
    const T* xin;
    T* xout;
    T* d;
    const T* b;
    const S* invd;
    T c1;
    T c2;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        cheb_body<OP, T, S>(blk, nblk, op, xin, xout, d, b, invd, c1, c2);
    }
};

// The first Chebyshev step from x = 0: d = c0 D^-1 b, x = d.
template <class T, class S = T>
__device__ __forceinline__ void cheb_first_body(unsigned blk, unsigned nblk, size_t n, const T* __restrict__ b, const S* __restrict__ invd,
                                                      T* __restrict__ x, T* __restrict__ d, T c0) {
    const size_t u = blk * (size_t)BLOCK + threadIdx.x;
    if (u >= n) return;
    const T v = c0 * ld<T>(invd, u) * b[u];
    d[u] = v;
    x[u] = v;
}
template <class T, class S = T>
__global__ void __launch_bounds__(BLOCK) k_cheb_first(size_t n, const T* __restrict__ b, const S* __restrict__ invd,
                                                      T* __restrict__ x, T* __restrict__ d, T c0) {
    cheb_first_body<T, S>(blockIdx.x, gridDim.x, n, b, invd, x, d, c0);
}
template <class T, class S = T>
struct KChebFirst {
    static constexpr int NT = BLOCK, MINB = 0;
    size_t n;
    const T* b;
    const S* invd;
    T* x;
    T* d;
    T c0;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        cheb_first_body<T, S>(blk, nblk, n, b, invd, x, d, c0);
    }
};

// invd = 1 / diag(A); 0 on fixed nodes.
template <class OP, class T>
__global__ void __launch_bounds__(BLOCK) k_inv_diag(OP op, T* __restrict__ invd) {
    const GridDims g = op.dims();
    const size_t node = xcd_block() * BLOCK + threadIdx.x;
    if (node >= g.nodes()) return;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    T dg[OP::C];
    op.diag(i, j, k, dg);
    const bool fx = is_fixed(g, i, j, k);
#pragma unroll
    for (int c = 0; c < OP::C; ++c) invd[OP::C * node + c] = fx ? T(0) : T(1) / dg[c];
}

// ------------------------------------------------------------------ grid transfers
// On each axis the fine grid has 2 n + 1 nodes (n coarse cells), or 2 n on an axis of an odd
// number of fine cells, whose last coarse node lies one fine cell past the box.
// Trilinear (hex8): fine node f gets coarse f/2 (even) or the mean of floor and ceil (odd),
// as a tensor product. Kuhn P1 (tet4, and tet10 -> tet4 on the same cells): fine node f is the
// midpoint of the Kuhn edge from floor(f/2) to ceil(f/2), componentwise.
template <class T, int C, bool KUHN_P1>
__device__ __forceinline__ void prolong_add_body(unsigned blk, unsigned nblk, GridDims fg, GridDims cg, const T* __restrict__ ec,
                                                       T* __restrict__ xf) {
    const size_t node = xcd_block(blk, nblk) * BLOCK + threadIdx.x;
    if (node >= fg.nodes()) return;
    int i, j, k;
    node_ijk(fg, node, i, j, k);
    if (is_fixed(fg, i, j, k)) return;
    T acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = T(0);
    const int lo[3] = {i >> 1, j >> 1, k >> 1};
    const int hi[3] = {(i + 1) >> 1, (j + 1) >> 1, (k + 1) >> 1};
    if (KUHN_P1) {
        const size_t a = lo[0] + (size_t)cg.nx * (lo[1] + (size_t)cg.ny * lo[2]);
        const size_t b = hi[0] + (size_t)cg.nx * (hi[1] + (size_t)cg.ny * hi[2]);
        const bool same = a == b;
        T va[C], vb[C];  // both ends loaded unconditionally (b = a for a coarse node)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            va[c] = ec[C * a + c];
            vb[c] = ec[C * b + c];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] = same ? va[c] : T(0.5) * (va[c] + vb[c]);
    } else {
        // Per axis: the lower coarse node with weight 1 (even coordinate) or 1/2 (odd), the upper
        // one with 0 or 1/2 (for an even coordinate it is the same node). All 8 corners are loaded
        // unconditionally, so that the loads are in flight together.
        const T w1[3] = {(i & 1) ? T(0.5) : T(0), (j & 1) ? T(0.5) : T(0), (k & 1) ? T(0.5) : T(0)};
        T v[8][C];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int cx = (q & 1) ? hi[0] : lo[0], cy = ((q >> 1) & 1) ? hi[1] : lo[1], cz = (q >> 2) ? hi[2] : lo[2];
            const size_t cn = cx + (size_t)cg.nx * (cy + (size_t)cg.ny * cz);
#pragma unroll
            for (int c = 0; c < C; ++c) v[q][c] = ec[C * cn + c];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const T w = ((q & 1) ? w1[0] : T(1) - w1[0]) * (((q >> 1) & 1) ? w1[1] : T(1) - w1[1]) *
                        ((q >> 2) ? w1[2] : T(1) - w1[2]);
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] += w * v[q][c];
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) xf[C * node + c] += acc[c];
}
template <class T, int C, bool KUHN_P1>
__global__ void __launch_bounds__(BLOCK) k_prolong_add(GridDims fg, GridDims cg, const T* __restrict__ ec,
                                                       T* __restrict__ xf) {
    prolong_add_body<T, C, KUHN_P1>(blockIdx.x, gridDim.x, fg, cg, ec, xf);
}
template <class T, int C, bool KUHN_P1>
struct KProlongAdd {
    static constexpr int NT = BLOCK, MINB = 0;
    GridDims fg;
    GridDims cg;
    const T* ec;
    T* xf;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        prolong_add_body<T, C, KUHN_P1>(blk, nblk, fg, cg, ec, xf);
    }
};

// The transpose: b_c = P^T r_f, gathered per coarse node; 0 on fixed coarse nodes. With invd,
// also the coarse level's first Chebyshev step from x = 0: d_c = x_c = c0 D^-1 b_c. With dense_of
// (the coarse level is the dense one), b_c goes to its dense index in dense_b instead of bc.
template <class T, int C, bool KUHN_P1, class S>
__device__ __forceinline__ void restrict_body(unsigned blk, unsigned nblk, GridDims fg, GridDims cg, const T* __restrict__ rf,
                                                    T* __restrict__ bc, const S* __restrict__ invd, T c0,
                                                    T* __restrict__ xc, T* __restrict__ dc,
                                                    const int32_t* __restrict__ dense_of, T* __restrict__ dense_b) {
    const size_t node = xcd_block(blk, nblk) * BLOCK + threadIdx.x;
    if (node >= cg.nodes()) return;
    int I, J, K;
    node_ijk(cg, node, I, J, K);
    T acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = T(0);
    if (!is_fixed(cg, I, J, K)) {
        const int fi = 2 * I, fj = 2 * J, fk = 2 * K;
        if (KUHN_P1) {
            // The coarse node itself, and the 14 midpoints of the Kuhn edges at it: directions
            // +-(1,0,0), +-(0,1,0), +-(0,0,1), +-(1,1,0), +-(1,0,1), +-(0,1,1), +-(1,1,1). All 15
            // values are loaded first (indices clamped to the grid, weight 0 outside), then summed
            // in that order.
            const int DIRS[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
            T v[15][C];
            bool in[15];
#pragma unroll
            for (int q = 0; q < 15; ++q) {
                const int s = (q - 1) / 2, sign = (q & 1) ? -1 : 1;
                const int a = q == 0 ? fi : fi + sign * DIRS[s][0], b = q == 0 ? fj : fj + sign * DIRS[s][1],
                          cz = q == 0 ? fk : fk + sign * DIRS[s][2];
                in[q] = a >= 0 && b >= 0 && cz >= 0 && a < fg.nx && b < fg.ny && cz < fg.nz;
                const size_t f = min(max(a, 0), fg.nx - 1) +
                                 (size_t)fg.nx * (min(max(b, 0), fg.ny - 1) + (size_t)fg.ny * min(max(cz, 0), fg.nz - 1));
#pragma unroll
                for (int c = 0; c < C; ++c) v[q][c] = rf[C * f + c];
            }
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] = in[0] ? v[0][c] : T(0);
#pragma unroll
            for (int q = 1; q < 15; ++q)
#pragma unroll
                for (int c = 0; c < C; ++c)
                    if (in[q]) acc[c] += T(0.5) * v[q][c];
        } else {
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int a = fi + dx, b = fj + dy, cz = fk + dz;
                        const bool v = a >= 0 && b >= 0 && cz >= 0 && a < fg.nx && b < fg.ny && cz < fg.nz;
                        const T w = v ? T(dx ? 0.5 : 1.0) * T(dy ? 0.5 : 1.0) * T(dz ? 0.5 : 1.0) : T(0);
                        const size_t f = min(max(a, 0), fg.nx - 1) +
                                         (size_t)fg.nx * (min(max(b, 0), fg.ny - 1) + (size_t)fg.ny * min(max(cz, 0), fg.nz - 1));
#pragma unroll
                        for (int c = 0; c < C; ++c) acc[c] += w * rf[C * f + c];
                    }
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const size_t u = C * node + c;
        if (dense_of) {
            const int32_t j = dense_of[u];
            if (j >= 0) dense_b[j] = acc[c];
            continue;
        }
        bc[u] = acc[c];
        if (invd) {  // the coarse level's first Chebyshev step (k_cheb_first, same arithmetic)
            const T v = c0 * ld<T>(invd, u) * acc[c];
            dc[u] = v;
            xc[u] = v;
        }
    }
}
template <class T, int C, bool KUHN_P1, class S>
__global__ void __launch_bounds__(BLOCK) k_restrict(GridDims fg, GridDims cg, const T* __restrict__ rf,
                                                    T* __restrict__ bc, const S* __restrict__ invd, T c0,
                                                    T* __restrict__ xc, T* __restrict__ dc,
                                                    const int32_t* __restrict__ dense_of, T* __restrict__ dense_b) {
    restrict_body<T, C, KUHN_P1, S>(blockIdx.x, gridDim.x, fg, cg, rf, bc, invd, c0, xc, dc, dense_of, dense_b);
}
template <class T, int C, bool KUHN_P1, class S>
struct KRestrict {
    static constexpr int NT = BLOCK, MINB = 0;
    GridDims fg;
    GridDims cg;
    const T* rf;
    T* bc;
    const S* invd;
    T c0;
    T* xc;
    T* dc;
    const int32_t* dense_of;
    T* dense_b;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        restrict_body<T, C, KUHN_P1, S>(blk, nblk, fg, cg, rf, bc, invd, c0, xc, dc, dense_of, dense_b);
    }
};

// ------------------------------------------------------------------ FP64 PCG pieces
// Scalars of one PCG solve, in device memory. The stopping test of the loop head
// (multigrid.rs:722-737) runs on the device: `active` says whether the next iteration runs (the
// relative residual above tol and fewer than it_limit iterations done), `go` is that decision
// taken at the head of the current iteration, and the kernels that change the solve's state (x,
// r, p and these scalars) do nothing when it is 0, so iterations queued past the end leave the
// solve as it was.
struct PcgScalars {
    double rz, pq, rr, bb, alpha, beta;
    double go, active, it, it_limit, tol;
};

// Configures a PCG pass: tolerance and the iterations it may take.
static __global__ void k_pcg_config(PcgScalars* s, double tol, double it_limit) {
    if (threadIdx.x != 0) return;
    s->tol = tol;
    s->it_limit = it_limit;
    s->it = 0.0;
    s->go = 0.0;
}

// The stopping test of the loop head on the current r.r (the host takes the same quotient).
__device__ __forceinline__ double pcg_next_active(const PcgScalars* s, double rr);

// what: 0 start (bb, rr, and the stopping test), 1 the head of an iteration (go = active, the
// count) then alpha = rz / pq, 2 rr (and the stopping test for the next iteration), 3 rz (first:
// rz = rz'), 4 beta = rz'/rz then rz = rz', 5 both 2 (from b) and 4 (from a). 1, 2, 4 and 5
// change nothing else in an iteration that does not run (go = 0). Run by one thread with the
// sums a (and b for what 0 and 5).
__device__ __forceinline__ void pcg_scalars_apply(PcgScalars* s, double a, double b, int what) {
    if (what == 1) {
        s->go = s->active;
        if (s->go != 0.0) s->it += 1.0;
    }
    if ((what == 1 || what == 2 || what == 4 || what == 5) && s->go == 0.0) return;
    switch (what) {
        case 0:
            s->bb = a;
            s->rr = b;
            s->active = pcg_next_active(s, b);
            break;
        case 1:
            s->pq = a;
            s->alpha = s->rz / a;
            break;
        case 2:
            s->rr = a;
            s->active = pcg_next_active(s, a);
            break;
        case 3:
            s->rz = a;
            break;
        case 4:
            s->beta = a / s->rz;
            s->rz = a;
            break;
        case 5:
            s->rr = b;
            s->active = pcg_next_active(s, b);
            s->beta = a / s->rz;
            s->rz = a;
            break;
    }
}

// The loop goes on while the residual is not at or below tol (multigrid.rs:721-727): a NaN
// residual goes on to the iteration limit, as the reference's comparison does.
__device__ __forceinline__ double pcg_next_active(const PcgScalars* s, double rr) {
    const double res = s->bb == 0.0 ? 0.0 : sqrt(rr) / sqrt(s->bb);
    return (!(res <= s->tol) && s->it < s->it_limit) ? 1.0 : 0.0;
}

// r = b - A x (FP64 operator); partial sums of b.b and r.r.
template <class OP>
__global__ void __launch_bounds__(BLOCK) k_pcg_start(OP op, const double* __restrict__ x, const double* __restrict__ b,
                                                     double* __restrict__ r, double* __restrict__ part_bb,
                                                     double* __restrict__ part_rr) {
    const GridDims g = op.dims();
    const size_t node = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    double bb = 0.0, rr = 0.0;
    if (node < g.nodes()) {
        int i, j, k;
        node_ijk(g, node, i, j, k);
        double out[OP::C];
        const bool fx = is_fixed(g, i, j, k);
        if (!fx) op.apply(x, i, j, k, out);
#pragma unroll
        for (int c = 0; c < OP::C; ++c) {
            const size_t u = OP::C * node + c;
            const double rv = fx ? 0.0 : b[u] - out[c];
            r[u] = rv;
            bb += b[u] * b[u];
            rr += rv * rv;
        }
    }
    bb = block_sum(bb);
    rr = block_sum(rr);
    if (threadIdx.x == 0) {
        part_bb[blockIdx.x] = bb;
        part_rr[blockIdx.x] = rr;
    }
}

// q = A p; partial sums of p.q.
template <class OP>
__device__ __forceinline__ void pcg_q_body(unsigned blk, unsigned nblk, OP op, const double* __restrict__ p, double* __restrict__ q,
                                                 double* __restrict__ part) {
    const GridDims g = op.dims();
    const size_t node = blk * (size_t)BLOCK + threadIdx.x;
    double pq = 0.0;
    if (node < g.nodes()) {
        int i, j, k;
        node_ijk(g, node, i, j, k);
        double out[OP::C];
        const bool fx = is_fixed(g, i, j, k);
        if (!fx) op.apply(p, i, j, k, out);
#pragma unroll
        for (int c = 0; c < OP::C; ++c) {
            const size_t u = OP::C * node + c;
            const double qv = fx ? 0.0 : out[c];
            q[u] = qv;
            pq += p[u] * qv;
        }
    }
    pq = block_sum(pq);
    if (threadIdx.x == 0) part[blk] = pq;
}
template <class OP>
__global__ void __launch_bounds__(BLOCK) k_pcg_q(OP op, const double* __restrict__ p, double* __restrict__ q,
                                                 double* __restrict__ part) {
    pcg_q_body<OP>(blockIdx.x, gridDim.x, op, p, q, part);
}
template <class OP>
struct KPcgQ {
    static constexpr int NT = BLOCK, MINB = 0;
    OP op;
    const double* p;
    double* q;
    double* part;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        pcg_q_body<OP>(blk, nblk, op, p, q, part);
    }
};

// r -= alpha q, rv = (V) r for the V-cycle; partial sums of r.r (grid-stride: any grid size). With invd, also the first
// Chebyshev step of the V-cycle's level 0 from x = 0: d0 = x0 = c0 D^-1 rv (the step k_cheb_first
// takes, in the same arithmetic). x += alpha p is left to k_pcg_dir, which reads p anyway.
template <class V, class S>
__device__ __forceinline__ void pcg_update_body(unsigned blk, unsigned nblk, size_t n, const PcgScalars* __restrict__ s,
                                                      double* __restrict__ r, const double* __restrict__ q,
                                                      V* __restrict__ rv, double* __restrict__ part,
                                                      const S* __restrict__ invd, V c0, V* __restrict__ x0,
                                                      V* __restrict__ d0) {
    double rr = 0.0;
    if (s->go != 0.0) {
        const double alpha = s->alpha;
        for (size_t u = blk * (size_t)BLOCK + threadIdx.x; u < n; u += (size_t)nblk * BLOCK) {
            const double ru = r[u] - alpha * q[u];
            r[u] = ru;
            const V bu = (V)ru;
            rv[u] = bu;
            if (invd) {
                const V v = c0 * ld<V>(invd, u) * bu;
                d0[u] = v;
                x0[u] = v;
            }
            rr += ru * ru;
        }
    }
    rr = block_sum(rr);
    if (threadIdx.x == 0) part[blk] = rr;
}
template <class V, class S>
__global__ void __launch_bounds__(BLOCK) k_pcg_update(size_t n, const PcgScalars* __restrict__ s,
                                                      double* __restrict__ r, const double* __restrict__ q,
                                                      V* __restrict__ rv, double* __restrict__ part,
                                                      const S* __restrict__ invd, V c0, V* __restrict__ x0,
                                                      V* __restrict__ d0) {
    pcg_update_body<V, S>(blockIdx.x, gridDim.x, n, s, r, q, rv, part, invd, c0, x0, d0);
}
template <class V, class S>
struct KPcgUpdate {
    static constexpr int NT = BLOCK, MINB = 0;
    size_t n;
    const PcgScalars* s;
    double* r;
    const double* q;
    V* rv;
    double* part;
    const S* invd;
    V c0;
    V* x0;
    V* d0;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        pcg_update_body<V, S>(blk, nblk, n, s, r, q, rv, part, invd, c0, x0, d0);
    }
};

// Partial sums of r.z (grid-stride), z read from zv (if zv; also stored to z as FP64 if z) or
// from z.
template <class V>
__device__ __forceinline__ void pcg_rz_body(unsigned blk, unsigned nblk, size_t n, const V* __restrict__ zv, double* __restrict__ z,
                                                  const double* __restrict__ r, double* __restrict__ part) {
    double rz = 0.0;
    for (size_t u = blk * (size_t)BLOCK + threadIdx.x; u < n; u += (size_t)nblk * BLOCK) {
        double zu;
        if (zv) {
            zu = (double)zv[u];
            if (z) z[u] = zu;
        } else {
            zu = z[u];
        }
        rz += r[u] * zu;
    }
    rz = block_sum(rz);
    if (threadIdx.x == 0) part[blk] = rz;
}
template <class V>
__global__ void __launch_bounds__(BLOCK) k_pcg_rz(size_t n, const V* __restrict__ zv, double* __restrict__ z,
                                                  const double* __restrict__ r, double* __restrict__ part) {
    pcg_rz_body<V>(blockIdx.x, gridDim.x, n, zv, z, r, part);
}
template <class V>
struct KPcgRz {
    static constexpr int NT = BLOCK, MINB = 0;
    size_t n;
    const V* zv;
    double* z;
    const double* r;
    double* part;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        pcg_rz_body<V>(blk, nblk, n, zv, z, r, part);
    }
};

// x += alpha p (the iteration's update of x, multigrid.rs:741), then p = z + beta p; first: p = z
// alone. z in the preconditioner's precision.
template <class V>
__device__ __forceinline__ void pcg_dir_body(unsigned blk, unsigned nblk, size_t n, const PcgScalars* __restrict__ s, const V* __restrict__ z,
                                                   double* __restrict__ p, double* __restrict__ x, int first) {
    const size_t u = blk * (size_t)BLOCK + threadIdx.x;
    if (u >= n) return;
    if (first) {
        p[u] = (double)z[u];
    } else if (s->go != 0.0) {
        const double pu = p[u];
        x[u] += s->alpha * pu;
        p[u] = (double)z[u] + s->beta * pu;
    }
}
template <class V>
__global__ void __launch_bounds__(BLOCK) k_pcg_dir(size_t n, const PcgScalars* __restrict__ s, const V* __restrict__ z,
                                                   double* __restrict__ p, double* __restrict__ x, int first) {
    pcg_dir_body<V>(blockIdx.x, gridDim.x, n, s, z, p, x, first);
}
template <class V>
struct KPcgDir {
    static constexpr int NT = BLOCK, MINB = 0;
    size_t n;
    const PcgScalars* s;
    const V* z;
    double* p;
    double* x;
    int first;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        pcg_dir_body<V>(blk, nblk, n, s, z, p, x, first);
    }
};

// Sum of `count` partials in a fixed order, by one block of RBLOCK threads.
__device__ __forceinline__ double sum_partials(const double* part, int count) {
    double v = 0.0;
    for (int t = threadIdx.x; t < count; t += RBLOCK) v += part[t];
    return block_sum_t<RBLOCK>(v);
}

// The scalars from block partials (pcg_scalars_apply): count_a of part_a, count_b of part_b.
__device__ __forceinline__ void pcg_scalars_body(unsigned blk, unsigned nblk, PcgScalars* s, const double* part_a, const double* part_b,
                                                       int count_a, int count_b, int what) {
    const double a = sum_partials(part_a, count_a);
    double b = 0.0;
    if (what == 0 || what == 5) b = sum_partials(part_b, count_b);
    if (threadIdx.x == 0) pcg_scalars_apply(s, a, b, what);
}
static __global__ void __launch_bounds__(RBLOCK) k_pcg_scalars(PcgScalars* s, const double* part_a, const double* part_b,
                                                       int count_a, int count_b, int what) {
    pcg_scalars_body(blockIdx.x, gridDim.x, s, part_a, part_b, count_a, count_b, what);
}
struct KPcgScalars {
    static constexpr int NT = RBLOCK, MINB = 0;
    PcgScalars* s;
    const double* part_a;
    const double* part_b;
    int count_a;
    int count_b;
    int what;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        pcg_scalars_body(blk, nblk, s, part_a, part_b, count_a, count_b, what);
    }
};

// ------------------------------------------------------------------ small vector kernels
template <class S, class D>
__global__ void __launch_bounds__(BLOCK) k_convert(size_t n, const S* __restrict__ a, D* __restrict__ b) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u < n) b[u] = (D)a[u];
}

template <class T>
__global__ void __launch_bounds__(BLOCK) k_fill(size_t n, T* __restrict__ a, T v) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u < n) a[u] = v;
}

// y = a x + b y.
static __global__ void __launch_bounds__(BLOCK) k_axpby(size_t n, double a, const double* __restrict__ x, double b,
                                                 double* __restrict__ y) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u < n) y[u] = a * x[u] + b * y[u];
}

// grid[map[k]] = src[k] (the rest of grid is left as it is).
static __global__ void __launch_bounds__(BLOCK) k_scatter(size_t n, const int64_t* __restrict__ map, const double* __restrict__ src,
                                                   double* __restrict__ grid) {
    const size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (k < n) grid[map[k]] = src[k];
}

// dst[k] = grid[map[k]].
static __global__ void __launch_bounds__(BLOCK) k_gather(size_t n, const int64_t* __restrict__ map, const double* __restrict__ grid,
                                                  double* __restrict__ dst) {
    const size_t k = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (k < n) dst[k] = grid[map[k]];
}

// Partial sums of a.b (FP64).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_dot(size_t n, const T* __restrict__ a, const T* __restrict__ b,
                                               double* __restrict__ part) {
    double v = 0.0;
    for (size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x; u < n; u += (size_t)gridDim.x * BLOCK)
        v += (double)a[u] * (double)b[u];
    v = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// out[0] = sum of the partials.
static __global__ void __launch_bounds__(RBLOCK) k_sum(const double* part, int count, double* out) {
    const double v = sum_partials(part, count);
    if (threadIdx.x == 0) out[0] = v;
}

// a = s * a with s = 1 / sqrt(out[0]) (power iteration normalisation).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_normalise(size_t n, T* __restrict__ a, const double* __restrict__ norm2) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u >= n) return;
    const double s = norm2[0] > 0.0 ? 1.0 / sqrt(norm2[0]) : 0.0;
    a[u] = (T)(s * (double)a[u]);
}

// Rounds an FP32 array to BF16: a keeps the rounded values (so that setup sees the operator the
// smoother applies) and h their 16-bit storage.
static __global__ void __launch_bounds__(BLOCK) k_round_bf16(size_t n, float* __restrict__ a, bf16_t* __restrict__ h) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u >= n) return;
    const bf16_t v = bf16_round(a[u]);
    h[u] = v;
    a[u] = bf16_value(v);
}

// y = D^-1 y (elementwise).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_scale(size_t n, T* __restrict__ y, const T* __restrict__ invd) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u < n) y[u] *= invd[u];
}

// A deterministic pseudo-random start vector on the unknowns (0 on fixed nodes via invd = 0).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_random(size_t n, T* __restrict__ a, const T* __restrict__ invd) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u >= n) return;
    uint64_t h = (u + 1) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    const double v = 0.5 + (double)(h >> 11) * (1.0 / 9007199254740992.0);
    a[u] = invd[u] != T(0) ? (T)v : T(0);
}

}  // namespace fem
