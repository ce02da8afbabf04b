// The tet10 elasticity operator on Kuhn cells (tet10.rs:90-120) fused with its per-node updates:
// y = A x, r = b - A x, one Chebyshev-Jacobi step, or the PCG product q = A p with p.q partials.
//
// A workgroup owns a tile of TX x TY x TZ cells, one thread per cell:
//  1. the tile's box of refined nodes (2 T + 1 per axis) of x is staged in LDS;
//  2. each thread computes the 6 Kuhn tetrahedra of its cell (tet10_element_kuhn) and sums their
//     30 forces each into the 27 refined positions of the cell, in registers;
//  3. three exchanges through LDS, along x, then y, then z, add the forces on the face p = 2 of a
//     cell to the face p = 0 of the next cell of the tile, so that every node of the box is then
//     held by exactly one thread with the tile's complete contribution;
//  4. each cell finishes in place the nodes at its positions p in {0, 1}^3 whose cells all lie in
//     the tile (the mode's update, with b, D^-1, d of all 8 loaded at once and x from LDS); every
//     other node it holds (at p = 2 along some axis, or on the tile's low face with another tile
//     below) goes out as the tile's partial sum to slot s of the node, bit d of s set when the
//     tile lies below the node along axis d.
// A second launch sums the slots of each such node (the nodes on the faces shared by tiles and on
// the grid's high faces) in slot order and finishes it. Every sum has a fixed order (no atomics),
// so the results are bitwise the same from run to run.
#pragma once

#include <vector>

#include "kernels.hpp"
#include "march.hpp"
#include "tet10_kuhn.hpp"

namespace fem {

enum { T10_APPLY = 0, T10_RESID = 1, T10_CHEB = 2, T10_PCGQ = 3 };

// The tile of the fused kernels per storage type: 256 threads in FP32, 128 in FP64 (whose cell
// state takes twice the registers). Long in x: a wave then spans whole runs of a grid row, so its
// node loads and stores touch fewer cache lines, and fewer tile faces cut across rows (measured on
// 72 x 36 x 36 cells against 8 x 8 x 4, 16 x 4 x 4 and 16 x 8 x 2: 25 us per FP32 Chebyshev step
// against 31, 27 and 27).
template <class T>
struct T10Tile;
template <>
struct T10Tile<float> {
    static constexpr int X = 32, Y = 4, Z = 2;
};
template <>
struct T10Tile<double> {
    static constexpr int X = 16, Y = 2, Z = 4;
};

template <class T, class S = T>
struct T10Args {
    GridDims g;          // the refined grid (2 c + 1 nodes per axis) and its fixed faces
    int cx, cy, cz;      // cells
    int tx, ty, tz;      // tiles per axis
    const S* coef;       // lambda at [t * cells + cell], mu at [(6 + t) * cells + cell], stored in S (T, or BF16)
    T ih0, ih1, ih2;     // inverse cell sizes
    T w;                 // quadrature weight: cell volume / 6 / 4
    const T* x;          // operand (p for PCGQ)
    T* y;                // y, r, x_out or q
    const T* b;
    T* d;
    const S* invd;       // D^-1, stored in S
    T c1, c2;
    double* part;        // PCGQ: block partials of x.y, the tile launch first, then the boundary launch
    T* P;                // partial sums: P[slot * 3 N + 3 node + c] over the N nodes, 8 slots
    const int* blist;    // the nodes finished by the boundary launch
    int nb;              // their count
    int nb1;             // blocks of the tile launch
};

// The role, along one axis, of refined position p (0..2) of the cell at tile index l of a tile
// whose first cell is c0, of extent tn, in a grid of cn cells: held = the cell holds this node
// after the exchanges; low / high = the node lies on the tile's low / high face and another tile
// lies beyond that face.
struct T10Role {
    bool held, low, high;
};
__host__ __device__ inline T10Role t10_role(int p, int l, int c0, int tn, int cn) {
    const int last = (cn - c0 < tn ? cn - c0 : tn) - 1;
    T10Role r;
    r.held = p < 2 || l == last;
    r.low = p == 0 && l == 0 && c0 > 0;
    r.high = p == 2 && l == last && c0 + tn < cn;
    return r;
}
// A refined coordinate on a face shared by two tiles of extent tn cells, in a grid of n nodes.
__host__ __device__ inline bool t10_on_face(int i, int tn, int n) { return i > 0 && i < n - 1 && i % (2 * tn) == 0; }

// The value at element idx of p when valid, else 0, without a branch (a load of element 0 in its
// place).
template <class T, class S>
__device__ __forceinline__ T t10_load(const S* p, bool valid, unsigned idx) {
    const T v = ld<T>(p, valid ? idx : 0u);
    return valid ? v : T(0);
}

// The 3 components of a node (unknowns u0 .. u0 + 2) when valid, else 0, branch-free; one base for
// the 3, so that they load as one vector access.
template <class T, class S>
__device__ __forceinline__ void t10_load3(const S* p, bool valid, unsigned u0, T v[3]) {
    const S* q = p + (valid ? u0 : 0u);
    const T a = ld<T>(q, 0), b = ld<T>(q, 1), c = ld<T>(q, 2);
    v[0] = valid ? a : T(0);
    v[1] = valid ? b : T(0);
    v[2] = valid ? c : T(0);
}

// The update of one node (unknowns u0 .. u0 + 2) of the mode from its operator value yv, its
// operand xin and (CHEB, RESID) its b, d and D^-1 already loaded; pq gathers x.q (PCGQ).
template <class T, int MODE, class S>
__device__ __forceinline__ void t10_update(const T10Args<T, S>& a, unsigned u0, bool fixed, const T yv[3], const T xin[3],
                                           const T bv[3], const T dv[3], const T iv[3], double& pq) {
    T out[3], dn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (MODE == T10_APPLY) {
            out[c] = fixed ? T(0) : yv[c];
        } else if (MODE == T10_RESID) {
            out[c] = fixed ? T(0) : bv[c] - yv[c];
        } else if (MODE == T10_CHEB) {
            dn[c] = fixed ? T(0) : a.c1 * dv[c] + a.c2 * iv[c] * (bv[c] - yv[c]);
            out[c] = fixed ? T(0) : xin[c] + dn[c];
        } else {
            out[c] = fixed ? T(0) : yv[c];
            pq += (double)xin[c] * (double)out[c];
        }
    }
    T* y = a.y + u0;
    y[0] = out[0];
    y[1] = out[1];
    y[2] = out[2];
    if (MODE == T10_CHEB) {
        T* d = a.d + u0;
        d[0] = dn[0];
        d[1] = dn[1];
        d[2] = dn[2];
    }
}

// The forces of the Kuhn tetrahedron of type TT of a cell, added to the cell's 27 positions F;
// base: the cell's first node in the LDS box (row stride RS values, LY rows per plane).
template <class T, int TT, int RS, int LY>
__device__ __forceinline__ void t10_cell_type(const T* base, const T ih[3], T lam, T mu, T w, T F[27][3]) {
    T u[30], yl[30];
#pragma unroll
    for (int n = 0; n < 10; ++n) {
        const int off = (t10_offset(TT, n, 2) * LY + t10_offset(TT, n, 1)) * RS + 3 * t10_offset(TT, n, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[3 * n + c] = base[off + c];
    }
    tet10_element_kuhn<T, TT>(u, ih, lam, mu, w, yl);
#pragma unroll
    for (int n = 0; n < 10; ++n) {
        const int pos = t10_offset(TT, n, 0) + 3 * t10_offset(TT, n, 1) + 9 * t10_offset(TT, n, 2);
#pragma unroll
        for (int c = 0; c < 3; ++c) F[pos][c] += yl[3 * n + c];
    }
}

// The position of face entry q (0..8) with coordinate p along axis AX.
template <int AX>
__host__ __device__ constexpr int t10_face_pos(int p, int q) {
    return AX == 0 ? p + 3 * (q % 3) + 9 * (q / 3) : AX == 1 ? (q % 3) + 3 * p + 9 * (q / 3) : (q % 3) + 3 * (q / 3) + 9 * p;
}

// Adds the face p = 2 along AX of each sending cell to the face p = 0 of the next cell along AX.
template <class T, int AX, int NEXT>
__device__ __forceinline__ void t10_exchange(T F[27][3], T* xch, int tid, bool send, bool recv) {
    if (send) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) xch[tid * 27 + 3 * q + c] = F[t10_face_pos<AX>(2, q)][c];
    }
    __syncthreads();
    if (recv) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) F[t10_face_pos<AX>(0, q)][c] += xch[(tid - NEXT) * 27 + 3 * q + c];
    }
    __syncthreads();
}

// FP32 asks for 2 waves per SIMD (at most 256 registers per lane), which two tiles of LDS per CU
// allow; FP64 takes 1. Index arithmetic is 32-bit: the host keeps every byte offset below 2^31.
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
__device__ __forceinline__ void t10_tile_body(unsigned blk, unsigned nblk, T10Args<T, S> a) {
    constexpr int NT = TX * TY * TZ, NW = NT / 64;
    constexpr int LX = 2 * TX + 1, LY = 2 * TY + 1, LZ = 2 * TZ + 1;
    constexpr int RS = 3 * LX;  // values per box row
    __shared__ T xs[LZ * LY * RS];
    __shared__ T xch[NT * 27];
    const int tid = threadIdx.x;
    const int tile = (int)xcd_block(blk, nblk);
    const int X0 = (tile % a.tx) * TX, Y0 = ((tile / a.tx) % a.ty) * TY, Z0 = (tile / (a.tx * a.ty)) * TZ;
    const unsigned NX = a.g.nx, NY = a.g.ny, NXY = NX * NY;
    const unsigned N3 = 3 * NXY * a.g.nz;
    const int lx = tid % TX, ly = (tid / TX) % TY, lz = tid / (TX * TY);
    const int ex = X0 + lx, ey = Y0 + ly, ez = Z0 + lz;
    const bool valid = ex < a.cx && ey < a.cy && ez < a.cz;
    // Every load of a phase is issued before the first of them is used, so that their latencies
    // overlap: the cell's 12 Lame coefficients, then the box. The loads are branch-free.
    const unsigned cells = (unsigned)a.cx * a.cy * a.cz;
    T lam[6], mu[6];
    {
        const unsigned cell = ex + (unsigned)a.cx * (ey + (unsigned)a.cy * ez);
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            lam[t] = t10_load<T>(a.coef, valid, t * cells + cell);
            mu[t] = t10_load<T>(a.coef, valid, (6 + t) * cells + cell);
        }
    }
    // 1. the box of x, zero beyond the grid: wave w takes rows w, w + NW, ...; a row is RS
    // contiguous values of the grid, in chunks of 64 lanes, read through a buffer resource (an
    // offset past the end reads 0).
    {
        const __amdgpu_buffer_rsrc_t rx = make_rsrc(a.x, N3);
        constexpr int ROWS = LY * LZ, RPW = (ROWS + NW - 1) / NW, CH = (RS + 63) / 64;
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
        const int I0 = 2 * X0, J0 = 2 * Y0, K0 = 2 * Z0;
        const int nI = min(LX, (int)NX - I0), nJ = min(LY, (int)NY - J0), nK = min(LZ, a.g.nz - K0);
        T v[RPW][CH];
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int r = wave + i * NW;  // wave-uniform
            const int by = r % LY, bz = r / LY;
            const bool row = r < ROWS && by < nJ && bz < nK;
            const unsigned base = 3 * (I0 + NX * (J0 + by) + NXY * (K0 + bz)) + lane;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int q = lane + 64 * c;
                v[i][c] = buffer_value<T>(rx, row && q < 3 * nI, base + 64 * c);
            }
        }
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int r = wave + i * NW;
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int q = lane + 64 * c;
                if (r < ROWS && q < RS) xs[r * RS + q] = v[i][c];
            }
        }
    }
    __syncthreads();
    // 2. the 6 tetrahedra of the cell
    T F[27][3];
#pragma unroll
    for (int p = 0; p < 27; ++p)
#pragma unroll
        for (int c = 0; c < 3; ++c) F[p][c] = T(0);
    const T* base = xs + (2 * lz * LY + 2 * ly) * RS + 3 * 2 * lx;
    if (valid) {
        const T ih[3] = {a.ih0, a.ih1, a.ih2};
        t10_cell_type<T, 0, RS, LY>(base, ih, lam[0], mu[0], a.w, F);
        t10_cell_type<T, 1, RS, LY>(base, ih, lam[1], mu[1], a.w, F);
        t10_cell_type<T, 2, RS, LY>(base, ih, lam[2], mu[2], a.w, F);
        t10_cell_type<T, 3, RS, LY>(base, ih, lam[3], mu[3], a.w, F);
        t10_cell_type<T, 4, RS, LY>(base, ih, lam[4], mu[4], a.w, F);
        t10_cell_type<T, 5, RS, LY>(base, ih, lam[5], mu[5], a.w, F);
    }
    // 3. the exchanges (every thread takes part: they hold barriers)
    const int lastx = min(TX, a.cx - X0) - 1, lasty = min(TY, a.cy - Y0) - 1, lastz = min(TZ, a.cz - Z0) - 1;
    t10_exchange<T, 0, 1>(F, xch, tid, lx < lastx, lx > 0 && lx <= lastx);
    t10_exchange<T, 1, TX>(F, xch, tid, ly < lasty, ly > 0 && ly <= lasty);
    t10_exchange<T, 2, TX * TY>(F, xch, tid, lz < lastz, lz > 0 && lz <= lastz);
    // 4. the nodes held by this cell: first the partial sums (p = 2 along some axis: held by the
    // last cells of the tile only), which frees most of F, then the 8 nodes p in {0, 1}^3, with
    // the loads of all 8 issued before the first update.
    double pq = 0.0;
    if (valid) {
        const unsigned node0 = 2 * ex + NX * (2 * ey) + NXY * (2 * ez);
#pragma unroll
        for (int pos = 0; pos < 27; ++pos) {
            const int px = pos % 3, py = (pos / 3) % 3, pz = pos / 9;
            if (px < 2 && py < 2 && pz < 2) continue;
            const T10Role rx = t10_role(px, lx, X0, TX, a.cx), ry = t10_role(py, ly, Y0, TY, a.cy),
                          rz = t10_role(pz, lz, Z0, TZ, a.cz);
            if (!(rx.held && ry.held && rz.held)) continue;
            const unsigned slot = (rx.high ? 1 : 0) | (ry.high ? 2 : 0) | (rz.high ? 4 : 0);
            T* q = a.P + (slot * N3 + 3 * (node0 + px + NX * py + NXY * pz));
            q[0] = F[pos][0];
            q[1] = F[pos][1];
            q[2] = F[pos][2];
        }
        constexpr int NV = MODE == T10_CHEB ? 3 : MODE == T10_RESID ? 1 : 0;
        T nv[3][8][3];
        bool fin[8];
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int px = p & 1, py = (p >> 1) & 1, pz = p >> 2;
            fin[p] = !(t10_role(px, lx, X0, TX, a.cx).low || t10_role(py, ly, Y0, TY, a.cy).low ||
                       t10_role(pz, lz, Z0, TZ, a.cz).low);
            const unsigned u0 = 3 * (node0 + px + NX * py + NXY * pz);
            if (NV > 0) t10_load3(a.b, fin[p], u0, nv[0][p]);
            if (NV > 1) t10_load3(a.d, fin[p], u0, nv[1][p]);
            if (NV > 2) t10_load3(a.invd, fin[p], u0, nv[2][p]);
        }
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int px = p & 1, py = (p >> 1) & 1, pz = p >> 2;
            const int pos = px + 3 * py + 9 * pz;
            const unsigned u0 = 3 * (node0 + px + NX * py + NXY * pz);
            if (!fin[p]) {  // on the tile's low face: slot 0
                T* q = a.P + u0;
                q[0] = F[pos][0];
                q[1] = F[pos][1];
                q[2] = F[pos][2];
                continue;
            }
            const T* xl = base + (pz * LY + py) * RS + 3 * px;
            const T xin[3] = {xl[0], xl[1], xl[2]};
            t10_update<T, MODE>(a, u0, is_fixed(a.g, 2 * ex + px, 2 * ey + py, 2 * ez + pz), F[pos], xin, nv[0][p],
                                nv[1][p], nv[2][p], pq);
        }
    }
    if (MODE == T10_PCGQ) {
        pq = block_sum_t<NT>(pq);
        if (tid == 0) a.part[blk] = pq;
    }
}
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
__global__ void __launch_bounds__(TX * TY * TZ, sizeof(T) == 4 ? 2 : 1) k_t10_tile(T10Args<T, S> a) {
    t10_tile_body<T, TX, TY, TZ, MODE, S>(blockIdx.x, gridDim.x, a);
}
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
struct KT10Tile {
    static constexpr int NT = TX * TY * TZ, MINB = (sizeof(T) == 4 ? 2 : 1);
    T10Args<T, S> a;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        t10_tile_body<T, TX, TY, TZ, MODE, S>(blk, nblk, a);
    }
};

// The boundary nodes: the sum of their slots in slot order, then the mode's update.
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
__device__ __forceinline__ void t10_bound_body(unsigned blk, unsigned nblk, T10Args<T, S> a) {
    const int e = blk * BLOCK + threadIdx.x;
    double pq = 0.0;
    if (e < a.nb) {
        const unsigned NX = a.g.nx, NY = a.g.ny, NXY = NX * NY;
        const unsigned N3 = 3 * NXY * a.g.nz;
        const unsigned node = (unsigned)a.blist[e];
        const int I = (int)(node % NX), J = (int)((node / NX) % NY), K = (int)(node / NXY);
        const int B = (t10_on_face(I, TX, NX) ? 1 : 0) | (t10_on_face(J, TY, NY) ? 2 : 0) | (t10_on_face(K, TZ, a.g.nz) ? 4 : 0);
        const unsigned u0 = 3 * node;
        T P[8][3], xin[3], bv[3] = {T(0), T(0), T(0)}, dv[3] = {T(0), T(0), T(0)}, iv[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 8; ++s) t10_load3(a.P, (s & ~B) == 0, s * N3 + u0, P[s]);
        t10_load3(a.x, true, u0, xin);
        if (MODE == T10_RESID || MODE == T10_CHEB) t10_load3(a.b, true, u0, bv);
        if (MODE == T10_CHEB) {
            t10_load3(a.d, true, u0, dv);
            t10_load3(a.invd, true, u0, iv);
        }
        T yv[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s & ~B) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) yv[c] += P[s][c];
        }
        t10_update<T, MODE>(a, u0, is_fixed(a.g, I, J, K), yv, xin, bv, dv, iv, pq);
    }
    if (MODE == T10_PCGQ) {
        pq = block_sum(pq);
        if (threadIdx.x == 0) a.part[a.nb1 + blk] = pq;
    }
}
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
__global__ void __launch_bounds__(BLOCK) k_t10_bound(T10Args<T, S> a) {
    t10_bound_body<T, TX, TY, TZ, MODE, S>(blockIdx.x, gridDim.x, a);
}
template <class T, int TX, int TY, int TZ, int MODE, class S = T>
struct KT10Bound {
    static constexpr int NT = BLOCK, MINB = 0;
    T10Args<T, S> a;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        t10_bound_body<T, TX, TY, TZ, MODE, S>(blk, nblk, a);
    }
};

// A refined coordinate that the boundary launch finishes: on a face shared by two tiles, or the
// grid's last.
__host__ __device__ inline bool t10_bound_coord(int i, int tn, int n) { return i == n - 1 || t10_on_face(i, tn, n); }

// The nodes that the boundary launch finishes, in increasing order, for a refined grid of c cells
// per axis under tiles of tn cells.
inline void t10_boundary(const int c[3], const int tn[3], std::vector<int>& list) {
    const int N[3] = {2 * c[0] + 1, 2 * c[1] + 1, 2 * c[2] + 1};
    list.clear();
    for (int k = 0; k < N[2]; ++k)
        for (int j = 0; j < N[1]; ++j)
            for (int i = 0; i < N[0]; ++i)
                if (t10_bound_coord(i, tn[0], N[0]) || t10_bound_coord(j, tn[1], N[1]) ||
                    t10_bound_coord(k, tn[2], N[2]))
                    list.push_back((int)(i + (size_t)N[0] * (j + (size_t)N[1] * k)));
}

}  // namespace fem
