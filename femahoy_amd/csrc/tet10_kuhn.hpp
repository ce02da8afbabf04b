// The tet10 elasticity element product on Kuhn cells (tet10.rs:90-120), with the Kuhn type fixed
// at compile time so that the element product runs with the type's gradient pattern written out,
// and the refined-grid offsets of the 10 nodes of each type. The fused operator kernels
// (tet10_fused.hpp) call it once per tetrahedron.
//
#pragma once

#include "kernels.hpp"
#include "march.hpp"
#include "ops.hpp"

namespace fem {


// The chain axes of Kuhn type t (the same order as KUHN) and the rank of axis d in the chain.
__host__ __device__ constexpr int kuhn_axis(int t, int r) {
    return t == 0 ? (r == 0 ? 0 : r == 1 ? 1 : 2)
         : t == 1 ? (r == 0 ? 0 : r == 1 ? 2 : 1)
         : t == 2 ? (r == 0 ? 1 : r == 1 ? 0 : 2)
         : t == 3 ? (r == 0 ? 1 : r == 1 ? 2 : 0)
         : t == 4 ? (r == 0 ? 2 : r == 1 ? 0 : 1)
                  : (r == 0 ? 2 : r == 1 ? 1 : 0);
}
__host__ __device__ constexpr int kuhn_rank(int t, int d) {
    return kuhn_axis(t, 0) == d ? 0 : kuhn_axis(t, 1) == d ? 1 : 2;
}
// Bit d of chain vertex v (0..3) of type t.
__host__ __device__ constexpr int kuhn_bit(int t, int v, int d) { return kuhn_rank(t, d) < v ? 1 : 0; }
// The refined offset along d of local node n (0..9) of type t within its cell (0..2).
__host__ __device__ constexpr int t10_offset(int t, int n, int d) {
    return n < 4 ? 2 * kuhn_bit(t, n, d)
         : n == 4 ? kuhn_bit(t, 0, d) + kuhn_bit(t, 1, d)
         : n == 5 ? kuhn_bit(t, 1, d) + kuhn_bit(t, 2, d)
         : n == 6 ? kuhn_bit(t, 0, d) + kuhn_bit(t, 2, d)
         : n == 7 ? kuhn_bit(t, 0, d) + kuhn_bit(t, 3, d)
         : n == 8 ? kuhn_bit(t, 1, d) + kuhn_bit(t, 3, d)
                  : kuhn_bit(t, 2, d) + kuhn_bit(t, 3, d);
}
// The local node of type t at refined cell offset (px, py, pz), or 15.
__host__ __device__ constexpr int t10_local_at(int t, int px, int py, int pz) {
    int r = 15;
    for (int n = 9; n >= 0; --n)
        if (t10_offset(t, n, 0) == px && t10_offset(t, n, 1) == py && t10_offset(t, n, 2) == pz) r = n;
    return r;
}
// The element product y = K u of one tet10 of Kuhn type TT with isotropic lambda, mu: the
// formulation of tet10_element with the barycentric gradients of the type written out. Along axis
// b, the gradients are -1/h_b at chain vertex m = rank(b), +1/h_b at m + 1 and 0 elsewhere, so
// every gradient product is a difference of two nodal values.
template <class T, int TT>
__host__ __device__ __forceinline__ void tet10_element_kuhn(const T u[30], const T ih[3], T lam, T mu, T w, T y[30]) {
    constexpr int NE[4][4] = {{-1, 4, 6, 7}, {4, -1, 5, 8}, {6, 5, -1, 9}, {7, 8, 9, -1}};
    constexpr int EDGE[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    const T qa = T(0.5854101966249685), qb = T(0.1381966011250105);  // tet10.rs:45-46
    T G[4][3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int m = kuhn_rank(TT, b);
        const int mm = NE[m][m + 1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T cm = u[3 * (m + 1) + a] - u[3 * m + a];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                T v;
                if (k == m) v = u[3 * mm + a] - u[3 * m + a];
                else if (k == m + 1) v = u[3 * (m + 1) + a] - u[3 * mm + a];
                else v = u[3 * NE[k][m + 1] + a] - u[3 * NE[k][m] + a];
                G[k][a][b] = ih[b] * (T(4) * v - cm);
            }
        }
    }
    T S[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) S[a][b] = G[0][a][b] + G[1][a][b] + G[2][a][b] + G[3][a][b];
    T sig[4][6];
    T ssum[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        T D[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) D[a][b] = qb * S[a][b] + (qa - qb) * G[q][a][b];
        const T tr = lam * (D[0][0] + D[1][1] + D[2][2]);
        sig[q][0] = tr + T(2) * mu * D[0][0];
        sig[q][1] = tr + T(2) * mu * D[1][1];
        sig[q][2] = tr + T(2) * mu * D[2][2];
        sig[q][3] = mu * (D[1][2] + D[2][1]);  // yz
        sig[q][4] = mu * (D[0][2] + D[2][0]);  // xz
        sig[q][5] = mu * (D[0][1] + D[1][0]);  // xy
#pragma unroll
        for (int s = 0; s < 6; ++s) ssum[s] += sig[q][s];
    }
    T Sp[4][6], S0[6];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < 6; ++s) Sp[k][s] = w * (qb * ssum[s] + (qa - qb) * sig[k][s]);
#pragma unroll
    for (int s = 0; s < 6; ++s) S0[s] = w * ssum[s];
    // the symmetric entry (a, b) of (xx, yy, zz, yz, xz, xy)
    auto sym = [](int a, int b) { return a == b ? a : 6 - a - b; };
    // corners: y_k = (4 Sp_k - S0) g_k
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int m = kuhn_rank(TT, b);
                const int s = sym(a, b);
                if (k == m + 1) acc += (T(4) * Sp[k][s] - S0[s]) * ih[b];
                else if (k == m) acc -= (T(4) * Sp[k][s] - S0[s]) * ih[b];
            }
            y[3 * k + a] = acc;
        }
    // edges (k, j): y = 4 (Sp_k g_j + Sp_j g_k)
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int k = EDGE[e][0], j = EDGE[e][1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int m = kuhn_rank(TT, b);
                const int s = sym(a, b);
                if (j == m + 1) acc += Sp[k][s] * ih[b];
                else if (j == m) acc -= Sp[k][s] * ih[b];
                if (k == m + 1) acc += Sp[j][s] * ih[b];
                else if (k == m) acc -= Sp[j][s] * ih[b];
            }
            y[3 * (4 + e) + a] = T(4) * acc;
        }
    }
}

}  // namespace fem
