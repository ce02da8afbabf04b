// Device-side operators of the structured box models, node-centric: each thread computes the
// whole row of one grid node, so no output is shared between threads (no atomics, bitwise
// reproducible results).
//
// Grid conventions: node (i, j, k) of a grid with nx x ny x nz nodes has index i + nx (j + ny k).
// A vector field stores the 3 components of a node side by side (unknown 3 node + c). A face of
// the box that is fixed (bit f of `fix`: 0 x=0, 1 x=max, 2 y=0, 3 y=max, 4 z=0, 5 z=max) holds
// no unknowns: every kernel writes 0 there, so the vectors are 0 on it and the operators read
// such neighbours as 0 without a test.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fem {

struct GridDims {
    int nx, ny, nz;  // nodes per axis
    uint32_t fix;    // fixed faces
    __host__ __device__ size_t nodes() const { return (size_t)nx * ny * nz; }
};

// BF16 storage: the upper 16 bits of an FP32 value, rounded to nearest even. Arrays of
// coefficients and Jacobi diagonals may be stored so (uint16_t) and widened to FP32 on load.
typedef uint16_t bf16_t;

__host__ __device__ inline bf16_t bf16_round(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(bf16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float v;
    __builtin_memcpy(&v, &u, 4);
    return v;
}

// Element i of an array stored in S, as T.
template <class T, class S>
__device__ __forceinline__ T ld(const S* p, size_t i) {
    if constexpr (sizeof(S) == 2) {
        return T(bf16_value(p[i]));
    } else {
        return T(p[i]);
    }
}

__device__ __forceinline__ bool is_fixed(const GridDims& g, int i, int j, int k) {
    return ((g.fix & 1u) && i == 0) || ((g.fix & 2u) && i == g.nx - 1) || ((g.fix & 4u) && j == 0) ||
           ((g.fix & 8u) && j == g.ny - 1) || ((g.fix & 16u) && k == 0) || ((g.fix & 32u) && k == g.nz - 1);
}

// ------------------------------------------------------------------ hex8 heat, scalar k per cell
// The 8 x 8 conduction matrix of a box cell (heat.rs:1246-1278) depends only on which axes the
// two corners differ along: entry (p, q) = k * c[(p xor q) as bits x|y<<1|z<<2]. The corner bits
// here are (x, y, z) in {0,1}^3, not the VTK order; the matrix is the same operator.
template <class T, class S = T>
struct Hex8K {
    GridDims g;
    const S* k;  // per cell, (nx-1)(ny-1)(nz-1), x fastest
    T c[8];

    __device__ __forceinline__ T cell_k(int ex, int ey, int ez) const {
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) return T(0);
        return ld<T>(k, ex + (size_t)(g.nx - 1) * (ey + (size_t)(g.ny - 1) * ez));
    }

    // Branch-free: addresses clamped into the grid, out-of-grid values selected to 0.
    __device__ __forceinline__ T apply(const T* x, int i, int j, int kk) const {
        int ia[3], ja[3], ka[3];
        bool va[3], vb[3], vc[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            va[t] = i + t - 1 >= 0 && i + t - 1 < g.nx;
            vb[t] = j + t - 1 >= 0 && j + t - 1 < g.ny;
            vc[t] = kk + t - 1 >= 0 && kk + t - 1 < g.nz;
            ia[t] = min(max(i + t - 1, 0), g.nx - 1);
            ja[t] = min(max(j + t - 1, 0), g.ny - 1);
            ka[t] = min(max(kk + t - 1, 0), g.nz - 1);
        }
        T xv[27];
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const T v = x[ia[dx] + (size_t)g.nx * (ja[dy] + (size_t)g.ny * ka[dz])];
                    xv[dx + 3 * dy + 9 * dz] = (va[dx] && vb[dy] && vc[dz]) ? v : T(0);
                }
        const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
        T kv[8];
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const int ex = i - 1 + cx, ey = j - 1 + cy, ez = kk - 1 + cz;
                    const bool v = ex >= 0 && ey >= 0 && ez >= 0 && ex < ncx && ey < ncy && ez < ncz;
                    const int cxx = min(max(ex, 0), max(ncx - 1, 0)), cyy = min(max(ey, 0), max(ncy - 1, 0)),
                              czz = min(max(ez, 0), max(ncz - 1, 0));
                    const T kval = ld<T>(k, cxx + (size_t)ncx * (cyy + (size_t)ncy * czz));
                    kv[cx + 2 * cy + 4 * cz] = v ? kval : T(0);
                }
        T acc = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const T ke = kv[cx + 2 * cy + 4 * cz];
                    const int lx = 1 - cx, ly = 1 - cy, lz = 1 - cz;
                    T s = T(0);
#pragma unroll
                    for (int qz = 0; qz < 2; ++qz)
#pragma unroll
                        for (int qy = 0; qy < 2; ++qy)
#pragma unroll
                            for (int qx = 0; qx < 2; ++qx)
                                s += c[(qx ^ lx) | ((qy ^ ly) << 1) | ((qz ^ lz) << 2)] *
                                     xv[(cx + qx) + 3 * (cy + qy) + 9 * (cz + qz)];
                    acc += ke * s;
                }
        return acc;
    }

    __device__ __forceinline__ T diag(int i, int j, int kk) const {
        T s = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) s += cell_k(i - 1 + cx, j - 1 + cy, kk - 1 + cz);
        return s * c[0];
    }
};

// ------------------------------------------------------------------ 27-point symmetric stencil
// The Galerkin coarse operators of the hex8 hierarchy. Coefficient s of node i couples i with
// i + OFF[s] (s = 0 the diagonal, s = 1..13 the forward half); the backward half is read from
// the neighbour's forward coefficient. Storage is s-major: a[s * nodes + i].
__constant__ static const int STENCIL_OFF[14][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {-1, -1, 1}, {0, -1, 1},
    {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},  {1, 0, 1}, {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};

template <class T>
struct Stencil27 {
    GridDims g;
    const T* a;

    // Branch-free: a neighbour outside the grid is read at the node itself and selected to 0. All
    // 27 coefficients and 27 values are loaded before the sum, so that the loads are in flight
    // together (the kernels on these levels are latency-bound).
    __device__ __forceinline__ T apply(const T* x, int i, int j, int k) const {
        const size_t n = g.nodes();
        const size_t me = i + (size_t)g.nx * (j + (size_t)g.ny * k);
        T af[14], xf[14], ab[14], xb[14];
        bool vf[14], vb[14];
        af[0] = a[me];
        xf[0] = x[me];
#pragma unroll
        for (int s = 1; s < 14; ++s) {
            const int dx = STENCIL_OFF[s][0], dy = STENCIL_OFF[s][1], dz = STENCIL_OFF[s][2];
            const long off = dx + (long)g.nx * (dy + (long)g.ny * dz);
            const int fi = i + dx, fj = j + dy, fk = k + dz;
            vf[s] = fi >= 0 && fj >= 0 && fk >= 0 && fi < g.nx && fj < g.ny && fk < g.nz;
            const int bi = i - dx, bj = j - dy, bk = k - dz;
            vb[s] = bi >= 0 && bj >= 0 && bk >= 0 && bi < g.nx && bj < g.ny && bk < g.nz;
            const size_t f = vf[s] ? me + off : me;
            const size_t b = vb[s] ? me - off : me;
            af[s] = a[s * n + me];
            xf[s] = x[f];
            ab[s] = a[s * n + b];
            xb[s] = x[b];
        }
        T acc = af[0] * xf[0];
#pragma unroll
        for (int s = 1; s < 14; ++s) acc += (vf[s] ? af[s] * xf[s] : T(0)) + (vb[s] ? ab[s] * xb[s] : T(0));
        return acc;
    }

    __device__ __forceinline__ T diag(int i, int j, int k) const {
        return a[i + (size_t)g.nx * (j + (size_t)g.ny * k)];
    }
};

// ------------------------------------------------------------------ Kuhn tetrahedra
// The reference's prism split of a structured box (prism.rs:144-154 on a footprint of quads cut
// at their lowest corner) gives, in each cell, the 6 tetrahedra of the Kuhn split: the chains
// 0 -> e_a -> e_a + e_b -> (1,1,1) for the 6 orders (a, b, c) of the axes. Type t has the axis
// order KUHN[t]. Its barycentric gradients are -e_a/h_a, e_a/h_a - e_b/h_b, e_b/h_b - e_c/h_c,
// e_c/h_c, and its volume hx hy hz / 6.
__constant__ static const int KUHN[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Bit d of chain vertex v (0..3) of Kuhn type t: set when axis d is among the first v axes of
// the chain KUHN[t].
__host__ __device__ constexpr int kuhn_chain_bit(int t, int v, int d) {
    const int a = t / 2 == 0 ? 0 : (t / 2 == 1 ? 1 : 2);
    const int b = t == 0 ? 1 : t == 1 ? 2 : t == 2 ? 0 : t == 3 ? 2 : t == 4 ? 0 : 1;
    const int rank = d == a ? 0 : (d == b ? 1 : 2);
    return rank < v ? 1 : 0;
}

// The vertex of a Kuhn tetrahedron of type t at cell corner (bx, by, bz), or -1.
__host__ __device__ constexpr int kuhn_vertex(int t, int bx, int by, int bz) {
    const int a = t / 2 == 0 ? 0 : (t / 2 == 1 ? 1 : 2);
    const int b = t == 0 ? 1 : t == 1 ? 2 : t == 2 ? 0 : t == 3 ? 2 : t == 4 ? 0 : 1;
    const int bits[3] = {bx, by, bz};
    const int n = bx + by + bz;
    if (n == 0) return 0;
    if (n == 3) return 3;
    if (n == 1) return bits[a] ? 1 : -1;
    return (bits[a] && bits[b]) ? 2 : -1;
}

// The gradient of barycentric coordinate v of a Kuhn tetrahedron of type t, with inverse cell
// sizes ih.
template <class T>
__device__ __forceinline__ void kuhn_grad(int t, int v, const T ih[3], T g[3]) {
    const int a = KUHN[t][0], b = KUHN[t][1], c = KUHN[t][2];
    g[0] = g[1] = g[2] = T(0);
    if (v == 0) {
        g[a] = -ih[a];
    } else if (v == 1) {
        g[a] = ih[a];
        g[b] = -ih[b];
    } else if (v == 2) {
        g[b] = ih[b];
        g[c] = -ih[c];
    } else {
        g[c] = ih[c];
    }
}

// Scalar P1 conduction on Kuhn tetrahedra (heat.rs:972-1000): k per tetrahedron, stored
// coef[t * cells + cell] for type t.
template <class T, class S = T>
struct Tet4Heat {
    GridDims g;
    const S* coef;
    T ih[3];
    T vol;

    // Node-centred, as Tet4Elastic::apply3: the 3 x 3 x 3 neighbourhood and the 8 cells'
    // coefficients first (clamped indices, zero coefficients outside the grid), then the 24
    // tetrahedra at the node in a fixed order.
    __device__ __forceinline__ T apply(const T* x, int i, int j, int kk) const {
        const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
        const size_t cells = (size_t)ncx * ncy * ncz;
        T u[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int ni = min(max(i + q % 3 - 1, 0), g.nx - 1), nj = min(max(j + (q / 3) % 3 - 1, 0), g.ny - 1),
                      nk = min(max(kk + q / 9 - 1, 0), g.nz - 1);
            u[q] = x[ni + (size_t)g.nx * (nj + (size_t)g.ny * nk)];
        }
        T kc[8][6];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ex = i - 1 + (c & 1), ey = j - 1 + ((c >> 1) & 1), ez = kk - 1 + (c >> 2);
            const bool in = ex >= 0 && ey >= 0 && ez >= 0 && ex < ncx && ey < ncy && ez < ncz;
            const size_t cell = min(max(ex, 0), ncx - 1) + (size_t)ncx * (min(max(ey, 0), ncy - 1) +
                                                                           (size_t)ncy * min(max(ez, 0), ncz - 1));
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const T k = ld<T>(coef, t * cells + cell);
                kc[c][t] = in ? k : T(0);
            }
        }
        T acc = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const int c = cx | (cy << 1) | (cz << 2);
                    const int lx = 1 - cx, ly = 1 - cy, lz = 1 - cz;
#pragma unroll
                    for (int t = 0; t < 6; ++t) {
                        const int me = kuhn_vertex(t, lx, ly, lz);
                        if (me < 0) continue;
                        // grad u = sum over the 4 vertices of u_v g_v
                        T gu[3] = {T(0), T(0), T(0)};
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int q = (cx + kuhn_chain_bit(t, v, 0)) +
                                          3 * ((cy + kuhn_chain_bit(t, v, 1)) + 3 * (cz + kuhn_chain_bit(t, v, 2)));
                            const T uv = u[q];
                            T gv[3];
                            kuhn_grad<T>(t, v, ih, gv);
                            gu[0] += uv * gv[0];
                            gu[1] += uv * gv[1];
                            gu[2] += uv * gv[2];
                        }
                        T gm[3];
                        kuhn_grad<T>(t, me, ih, gm);
                        acc += kc[c][t] * vol * (gm[0] * gu[0] + gm[1] * gu[1] + gm[2] * gu[2]);
                    }
                }
        return acc;
    }

    __device__ __forceinline__ T diag(int i, int j, int kk) const {
        const size_t cells = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
        T acc = T(0);
        for (int cz = 0; cz < 2; ++cz)
            for (int cy = 0; cy < 2; ++cy)
                for (int cx = 0; cx < 2; ++cx) {
                    const int ex = i - 1 + cx, ey = j - 1 + cy, ez = kk - 1 + cz;
                    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
                    const size_t cell = ex + (size_t)(g.nx - 1) * (ey + (size_t)(g.ny - 1) * ez);
                    for (int t = 0; t < 6; ++t) {
                        const int me = kuhn_vertex(t, 1 - cx, 1 - cy, 1 - cz);
                        if (me < 0) continue;
                        T gm[3];
                        kuhn_grad<T>(t, me, ih, gm);
                        acc += ld<T>(coef, t * cells + cell) * vol * (gm[0] * gm[0] + gm[1] * gm[1] + gm[2] * gm[2]);
                    }
                }
        return acc;
    }
};

// Isotropic linear elasticity on Kuhn tetrahedra, constant strain (the operator of
// tet10.rs:90-120 with constant gradients): lambda and mu per tetrahedron, stored
// coef[(0 * 6 + t) * cells + cell] (lambda) and coef[(1 * 6 + t) * cells + cell] (mu).
template <class T, class S = T>
struct Tet4Elastic {
    GridDims g;
    const S* coef;
    T ih[3];
    T vol;

    // Node-centred: the 3 x 3 x 3 node neighbourhood and the coefficients of the 8 cells around
    // the node are loaded first, all at once (indices clamped to the grid; a cell outside it gets
    // zero coefficients), then the 24 tetrahedra at the node are summed in a fixed order.
    __device__ __forceinline__ void apply3(const T* x, int i, int j, int kk, T out[3]) const {
        const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
        const size_t cells = (size_t)ncx * ncy * ncz;
        T u[27][3];
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int ni = min(max(i + q % 3 - 1, 0), g.nx - 1), nj = min(max(j + (q / 3) % 3 - 1, 0), g.ny - 1),
                      nk = min(max(kk + q / 9 - 1, 0), g.nz - 1);
            const size_t node = ni + (size_t)g.nx * (nj + (size_t)g.ny * nk);
#pragma unroll
            for (int a = 0; a < 3; ++a) u[q][a] = x[3 * node + a];
        }
        T lm[8][6][2];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ex = i - 1 + (c & 1), ey = j - 1 + ((c >> 1) & 1), ez = kk - 1 + (c >> 2);
            const bool in = ex >= 0 && ey >= 0 && ez >= 0 && ex < ncx && ey < ncy && ez < ncz;
            const size_t cell = min(max(ex, 0), ncx - 1) + (size_t)ncx * (min(max(ey, 0), ncy - 1) +
                                                                           (size_t)ncy * min(max(ez, 0), ncz - 1));
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const T lam = ld<T>(coef, (0 * 6 + t) * cells + cell), mu = ld<T>(coef, (1 * 6 + t) * cells + cell);
                lm[c][t][0] = in ? lam : T(0);
                lm[c][t][1] = in ? mu : T(0);
            }
        }
        out[0] = out[1] = out[2] = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const int c = cx | (cy << 1) | (cz << 2);
                    const int lx = 1 - cx, ly = 1 - cy, lz = 1 - cz;
#pragma unroll
                    for (int t = 0; t < 6; ++t) {
                        const int me = kuhn_vertex(t, lx, ly, lz);
                        if (me < 0) continue;
                        // G[a][b] = d u_a / d x_b = sum_v u_va g_v[b]
                        T G[3][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            // chain vertex v of type t: the first v axes of the chain set
                            const int bx = kuhn_chain_bit(t, v, 0), by = kuhn_chain_bit(t, v, 1), bz = kuhn_chain_bit(t, v, 2);
                            const int q = (cx + bx) + 3 * ((cy + by) + 3 * (cz + bz));
                            T gv[3];
                            kuhn_grad<T>(t, v, ih, gv);
#pragma unroll
                            for (int a = 0; a < 3; ++a)
#pragma unroll
                                for (int b = 0; b < 3; ++b) G[a][b] += u[q][a] * gv[b];
                        }
                        const T lam = lm[c][t][0], mu = lm[c][t][1];
                        const T tr = G[0][0] + G[1][1] + G[2][2];
                        T gm[3];
                        kuhn_grad<T>(t, me, ih, gm);
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            T s = T(0);
#pragma unroll
                            for (int b = 0; b < 3; ++b) {
                                const T sig = (a == b ? lam * tr : T(0)) + mu * (G[a][b] + G[b][a]);
                                s += sig * gm[b];
                            }
                            out[a] += vol * s;
                        }
                    }
                }
    }

    __device__ __forceinline__ void diag3(int i, int j, int kk, T out[3]) const {
        const size_t cells = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
        out[0] = out[1] = out[2] = T(0);
        for (int cz = 0; cz < 2; ++cz)
            for (int cy = 0; cy < 2; ++cy)
                for (int cx = 0; cx < 2; ++cx) {
                    const int ex = i - 1 + cx, ey = j - 1 + cy, ez = kk - 1 + cz;
                    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx - 1 || ey >= g.ny - 1 || ez >= g.nz - 1) continue;
                    const size_t cell = ex + (size_t)(g.nx - 1) * (ey + (size_t)(g.ny - 1) * ez);
                    for (int t = 0; t < 6; ++t) {
                        const int me = kuhn_vertex(t, 1 - cx, 1 - cy, 1 - cz);
                        if (me < 0) continue;
                        T gm[3];
                        kuhn_grad<T>(t, me, ih, gm);
                        const T lam = ld<T>(coef, (0 * 6 + t) * cells + cell);
                        const T mu = ld<T>(coef, (1 * 6 + t) * cells + cell);
                        const T g2 = gm[0] * gm[0] + gm[1] * gm[1] + gm[2] * gm[2];
                        for (int a = 0; a < 3; ++a) out[a] += vol * ((lam + mu) * gm[a] * gm[a] + mu * g2);
                    }
                }
    }
};

// ------------------------------------------------------------------ tet10 on Kuhn cells
// The 10 nodes of a Kuhn tetrahedron sit on the grid refined twice: its corners at 2 v and its
// edge points at v_a + v_b, with the chain vertices v_0..v_3 (in cell bits). Local nodes 0..3
// are the corners, 4..9 the edges (0,1), (1,2), (0,2), (0,3), (1,3), (2,3) (tet10.rs:38).
//
// The operator itself runs in the fused kernels of tet10_fused.hpp; Tet10Op below gives the
// diagonal of its rows (setup).
__constant__ static const int TET10_EDGE[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

// Chain vertex v (0..3) of Kuhn type t, as cell bits.
__host__ __device__ inline void kuhn_chain(int t, int v, int bits[3]) {
    const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    bits[0] = bits[1] = bits[2] = 0;
    for (int s = 0; s < v; ++s) bits[order[t][s]] = 1;
}

// The local node (0..9) of Kuhn type t at refined cell position (px, py, pz) in {0,1,2}^3, or -1.
__host__ __device__ inline int tet10_local(int t, int px, int py, int pz) {
    int vb[4][3];
    for (int v = 0; v < 4; ++v) kuhn_chain(t, v, vb[v]);
    for (int v = 0; v < 4; ++v)
        if (2 * vb[v][0] == px && 2 * vb[v][1] == py && 2 * vb[v][2] == pz) return v;
    const int E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    for (int e = 0; e < 6; ++e) {
        const int* a = vb[E[e][0]];
        const int* b = vb[E[e][1]];
        if (a[0] + b[0] == px && a[1] + b[1] == py && a[2] + b[2] == pz) return 4 + e;
    }
    return -1;
}

// The diagonal of the tet10 operator's rows, from the element rule of tet10.rs:90-120 (setup).
// coef: lambda at [t * cells + cell], mu at [(6 + t) * cells + cell] (cells of the coarse grid).
template <class T>
struct Tet10Op {
    GridDims g;  // refined grid: 2 n + 1 nodes per axis
    const T* coef;
    T ih[3];
    T vol;

    // The diagonal, recomputed from the element rule (setup only).
    __device__ void diag3(int i, int j, int k, T out[3]) const {
        const int cx = (g.nx - 1) / 2, cy = (g.ny - 1) / 2, cz = (g.nz - 1) / 2;
        const size_t cells = (size_t)cx * cy * cz;
        out[0] = out[1] = out[2] = T(0);
        const T qa = T(0.5854101966249685), qb = T(0.1381966011250105);
        // the cells whose refined positions 0..2 hold the node along each axis
        for (int ez = max(0, (k - 1) / 2); ez <= min(cz - 1, k / 2); ++ez) {
            if (k < 2 * ez || k > 2 * ez + 2) continue;
            for (int ey = max(0, (j - 1) / 2); ey <= min(cy - 1, j / 2); ++ey) {
                if (j < 2 * ey || j > 2 * ey + 2) continue;
                for (int ex = max(0, (i - 1) / 2); ex <= min(cx - 1, i / 2); ++ex) {
                    if (i < 2 * ex || i > 2 * ex + 2) continue;
                    const size_t cell = ex + (size_t)cx * (ey + (size_t)cy * ez);
                    for (int t = 0; t < 6; ++t) {
                        const int loc = tet10_local(t, i - 2 * ex, j - 2 * ey, k - 2 * ez);
                        if (loc < 0) continue;
                        T gg[4][3];
                        for (int v = 0; v < 4; ++v) kuhn_grad<T>(t, v, ih, gg[v]);
                        const T lam = coef[t * cells + cell], mu = coef[(6 + t) * cells + cell];
                        for (int q = 0; q < 4; ++q) {
                            T L[4];
                            for (int v = 0; v < 4; ++v) L[v] = v == q ? qa : qb;
                            T dn[3];
                            if (loc < 4) {
                                for (int a = 0; a < 3; ++a) dn[a] = (T(4) * L[loc] - T(1)) * gg[loc][a];
                            } else {
                                const int p = TET10_EDGE[loc - 4][0], r = TET10_EDGE[loc - 4][1];
                                for (int a = 0; a < 3; ++a) dn[a] = T(4) * (L[p] * gg[r][a] + L[r] * gg[p][a]);
                            }
                            const T n2 = dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2];
                            for (int a = 0; a < 3; ++a) out[a] += vol / T(4) * ((lam + mu) * dn[a] * dn[a] + mu * n2);
                        }
                    }
                }
            }
        }
    }
};

}  // namespace fem
