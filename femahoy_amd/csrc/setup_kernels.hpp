// Setup kernels (FP64): Galerkin coarse operators of the hex8 hierarchy, coefficient averaging of
// the nested Kuhn hierarchy, and the probing that assembles a dense coarse matrix.
//
// A coarse level has ceil(n / 2) cells on an axis of n fine cells. On an odd axis the last coarse
// cell reaches one fine cell past the box: its children outside the box are void (no material),
// so they add nothing to a Galerkin product or a coefficient sum. The coarse functions stay
// nested (a fine cell is a child of the 2x refinement of its coarse cell), and every coarse cell
// keeps at least its corner child, so the coarse operators stay positive definite.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fem {

// The trilinear interpolation from the 8 corners of a coarse cell to the 8 corners of its child
// c (bits x|y<<1|z<<2): P[c][p * 8 + q] for fine corner p and coarse corner q.
__constant__ static double HEX_CHILD_P[8][64];

// Level 1 of the hex8 hierarchy: K_C = sum over the 8 children c of k_c M_c, M_c = P_c^T K1 P_c
// (K1 the fine cell matrix for k = 1), the Galerkin product restricted to the coarse cell
// (exact: a child's nodes interpolate from its coarse cell's corners only).
__global__ void k_hex_cells_from_k(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz, const double* __restrict__ kf,
                                   const double* __restrict__ M, double* __restrict__ Kc) {
    const size_t cell = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    const size_t nc = (size_t)ncx * ncy * ncz;
    if (cell >= nc) return;
    const int I = (int)(cell % ncx), J = (int)((cell / ncx) % ncy), K = (int)(cell / ((size_t)ncx * ncy));
    double acc[64];
    for (int e = 0; e < 64; ++e) acc[e] = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int fx = 2 * I + (c & 1), fy = 2 * J + ((c >> 1) & 1), fz = 2 * K + (c >> 2);
        if (fx >= nfx || fy >= nfy || fz >= nfz) continue;
        const double k = kf[fx + (size_t)nfx * (fy + (size_t)nfy * fz)];
        for (int e = 0; e < 64; ++e) acc[e] += k * M[c * 64 + e];
    }
    for (int e = 0; e < 64; ++e) Kc[cell * 64 + e] = acc[e];
}

// A coarse hex8 level rediscretized: the mean conductivity of the 8 children of each coarse cell.
__global__ void k_hex_average(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz, const double* __restrict__ kf,
                              double* __restrict__ kc) {
    const size_t cell = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    const size_t nc = (size_t)ncx * ncy * ncz;
    if (cell >= nc) return;
    const int I = (int)(cell % ncx), J = (int)((cell / ncx) % ncy), K = (int)(cell / ((size_t)ncx * ncy));
    double s = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int fx = 2 * I + (c & 1), fy = 2 * J + ((c >> 1) & 1), fz = 2 * K + (c >> 2);
        if (fx >= nfx || fy >= nfy || fz >= nfz) continue;
        s += kf[fx + (size_t)nfx * (fy + (size_t)nfy * fz)];
    }
    kc[cell] = s / 8.0;
}

// Level l+1 from level l cell matrices: K_C = sum_c P_c^T K_c P_c.
__global__ void k_hex_cells_from_cells(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz,
                                       const double* __restrict__ Kf, double* __restrict__ Kc) {
    const size_t cell = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    const size_t nc = (size_t)ncx * ncy * ncz;
    if (cell >= nc) return;
    const int I = (int)(cell % ncx), J = (int)((cell / ncx) % ncy), K = (int)(cell / ((size_t)ncx * ncy));
    double acc[64];
    for (int e = 0; e < 64; ++e) acc[e] = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int fx = 2 * I + (c & 1), fy = 2 * J + ((c >> 1) & 1), fz = 2 * K + (c >> 2);
        if (fx >= nfx || fy >= nfy || fz >= nfz) continue;
        const double* A = Kf + (fx + (size_t)nfx * (fy + (size_t)nfy * fz)) * 64;
        const double* P = HEX_CHILD_P[c];
        // T = A P (8 x 8), then acc += P^T T
        double T[64];
        for (int p = 0; p < 8; ++p)
            for (int q = 0; q < 8; ++q) {
                double s = 0.0;
                for (int r = 0; r < 8; ++r) s += A[p * 8 + r] * P[r * 8 + q];
                T[p * 8 + q] = s;
            }
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                double s = 0.0;
                for (int p = 0; p < 8; ++p) s += P[p * 8 + a] * T[p * 8 + b];
                acc[a * 8 + b] += s;
            }
    }
    for (int e = 0; e < 64; ++e) Kc[cell * 64 + e] = acc[e];
}

// The 27-point stencil of the assembled cell matrices: a[s * nodes + i] couples node i with
// i + STENCIL_OFF[s].
__global__ void k_stencil_from_cells(GridDims g, const double* __restrict__ Kc, double* __restrict__ a) {
    const size_t node = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    const size_t n = g.nodes();
    if (node >= n) return;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
    for (int s = 0; s < 14; ++s) {
        const int dx = STENCIL_OFF[s][0], dy = STENCIL_OFF[s][1], dz = STENCIL_OFF[s][2];
        const int oi = i + dx, oj = j + dy, ok = k + dz;
        double v = 0.0;
        if (oi >= 0 && oj >= 0 && ok >= 0 && oi < g.nx && oj < g.ny && ok < g.nz) {
            for (int cz = 0; cz < 2; ++cz)
                for (int cy = 0; cy < 2; ++cy)
                    for (int cx = 0; cx < 2; ++cx) {
                        const int ex = i - 1 + cx, ey = j - 1 + cy, ez = k - 1 + cz;
                        if (ex < 0 || ey < 0 || ez < 0 || ex >= ncx || ey >= ncy || ez >= ncz) continue;
                        const int bx = oi - ex, by = oj - ey, bz = ok - ez;
                        if (bx < 0 || by < 0 || bz < 0 || bx > 1 || by > 1 || bz > 1) continue;
                        const int li = (i - ex) | ((j - ey) << 1) | ((k - ez) << 2);
                        const int lj = bx | (by << 1) | (bz << 2);
                        v += Kc[(ex + (size_t)ncx * (ey + (size_t)ncy * ez)) * 64 + li * 8 + lj];
                    }
        }
        a[s * n + node] = v;
    }
}

// Coarse Kuhn coefficients: the mean over the 8 fine tetrahedra inside each coarse one (the
// Galerkin product of nested P1 spaces: a coarse gradient is constant on the coarse tetrahedron,
// so P^T A P integrates the fine material). kids[t][m] = (child cell bits) * 8 + fine type.
__constant__ static int KUHN_KIDS[6][8];

__global__ void k_kuhn_average(int ncx, int ncy, int ncz, int nfx, int nfy, int nfz, int q,
                               const double* __restrict__ fine, double* __restrict__ coarse) {
    const size_t nc = (size_t)ncx * ncy * ncz;
    const size_t tid = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (tid >= (size_t)q * 6 * nc) return;
    const size_t cell = tid % nc;
    const int t = (int)((tid / nc) % 6), comp = (int)(tid / (6 * nc));
    const int I = (int)(cell % ncx), J = (int)((cell / ncx) % ncy), K = (int)(cell / ((size_t)ncx * ncy));
    const size_t nf = (size_t)nfx * nfy * nfz;
    double s = 0.0;
    for (int m = 0; m < 8; ++m) {
        const int code = KUHN_KIDS[t][m];
        const int c = code / 8, ft = code % 8;
        const int fx = 2 * I + (c & 1), fy = 2 * J + ((c >> 1) & 1), fz = 2 * K + (c >> 2);
        if (fx >= nfx || fy >= nfy || fz >= nfz) continue;
        s += fine[((size_t)comp * 6 + ft) * nf + fx + (size_t)nfx * (fy + (size_t)nfy * fz)];
    }
    coarse[((size_t)comp * 6 + t) * nc + cell] = s / 8.0;
}

// Probing of a local operator: unit values on the unknowns of component `comp` at the nodes of
// one residue class mod S per axis (S = 2 R + 1 for a coupling radius R), so that each row sees at
// most one of them.
__global__ void k_probe_fill(GridDims g, int C, int S, int c0, int c1, int c2, int comp,
                             const int64_t* __restrict__ dense_of, double* __restrict__ x) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u >= g.nodes() * C) return;
    const size_t node = u / C;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    const bool hit = (int)(u % C) == comp && i % S == c0 && j % S == c1 && k % S == c2 && dense_of[u] >= 0;
    x[u] = hit ? 1.0 : 0.0;
}

__global__ void k_probe_extract(GridDims g, int C, int S, int c0, int c1, int c2, int comp,
                                const int64_t* __restrict__ dense_of, const double* __restrict__ y,
                                double* __restrict__ A, int64_t n) {
    const size_t u = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    if (u >= g.nodes() * C) return;
    const int64_t row = dense_of[u];
    if (row < 0) return;
    const size_t node = u / C;
    int i, j, k;
    node_ijk(g, node, i, j, k);
    const int R = S / 2;
    auto near = [&](int p, int col) {
        int off = ((col - p) % S + S) % S;
        if (off > R) off -= S;
        return p + off;
    };
    const int a = near(i, c0), b = near(j, c1), c = near(k, c2);
    if (a < 0 || b < 0 || c < 0 || a >= g.nx || b >= g.ny || c >= g.nz) return;
    const int64_t col = dense_of[(a + (size_t)g.nx * (b + (size_t)g.ny * c)) * C + comp];
    if (col < 0) return;
    A[row * n + col] = y[u];
}

// The probe's values into a CSR of the free unknowns (row r: grid unknown free_grid[r]): every
// entry whose column (a free unknown) has the probe's colour (residues c0, c1, c2 mod S of its
// node, component comp) takes y at the row's grid unknown. Within the coupling radius (S = 2R + 1)
// a row meets one node of each colour, so each entry is written by exactly one probe.
__global__ void k_probe_csr(uint64_t rows, GridDims g, int C, int S, int c0, int c1, int c2, int comp,
                            const uint64_t* __restrict__ ptr, const uint32_t* __restrict__ idx,
                            const int64_t* __restrict__ free_grid, const double* __restrict__ y,
                            double* __restrict__ val) {
    const uint64_t r = blockIdx.x * (uint64_t)BLOCK + threadIdx.x;
    if (r >= rows) return;
    const double yr = y[free_grid[r]];
    for (uint64_t k = ptr[r]; k < ptr[r + 1]; ++k) {
        const int64_t gu = free_grid[idx[k]];
        const int64_t node = gu / C;
        int i, j, kk;
        node_ijk(g, (size_t)node, i, j, kk);
        if ((int)(gu % C) == comp && i % S == c0 && j % S == c1 && kk % S == c2) val[k] = yr;
    }
}

}  // namespace fem
