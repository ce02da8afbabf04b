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
// The element product y = K u of one tet10 of Kuhn type TT with isotropic lambda, mu (the matrix
// of tet10.rs:90-120: the 4-point rule of degree 2 is exact here), in vertex form. On an affine
// tetrahedron grad u is linear, so the stress is linear with vertex values s_m = sigma(G_m), G_m
// the gradient at vertex m; grad N_i is linear too, and int L_m L_k = V (1 + [m = k]) / 20 gives
// y_i = sum_m grad N_i(v_m) : t_m with t_m = V / 20 (s_m + sum_k s_k). With the barycentric
// gradients of a Kuhn chain (along axis b: -1/h_b at vertex m = rank(b), +1/h_b at m + 1, 0
// elsewhere), every gradient product is a difference of two values:
//  - G_m[a][b] = (4 v - c) / h_b, v and c differences of nodal values of component a;
//  - corner k: y_k = (4 t_k - sum_m t_m) grad L_k, with sum_m t_m = 5 V / 20 sum_k s_k;
//  - edge (i, j): y_ij = 4 (t_i grad L_j + t_j grad L_i).
// w is the cell volume / 24 (V / 4 of each of the 6 tetrahedra).
template <class T, int TT>
__host__ __device__ __forceinline__ void tet10_element_kuhn(const T u[30], const T ih[3], T lam, T mu, T w, T y[30]) {
    constexpr int NE[4][4] = {{-1, 4, 6, 7}, {4, -1, 5, 8}, {6, 5, -1, 9}, {7, 8, 9, -1}};
    constexpr int EDGE[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    // the vertex gradients G[k][a][b] = v * (4 / h_b) - c / h_b
    T G[4][3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int m = kuhn_rank(TT, b);
        const int mm = NE[m][m + 1];
        const T i4 = T(4) * ih[b];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T ci = (u[3 * (m + 1) + a] - u[3 * m + a]) * ih[b];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                T v;
                if (k == m) v = u[3 * mm + a] - u[3 * m + a];
                else if (k == m + 1) v = u[3 * (m + 1) + a] - u[3 * mm + a];
                else v = u[3 * NE[k][m + 1] + a] - u[3 * NE[k][m] + a];
                G[k][a][b] = v * i4 - ci;
            }
        }
    }
    // vertex stresses with lambda, mu scaled by V / 20, in (xx, yy, zz, yz, xz, xy)
    const T v20 = w * T(0.2);
    const T l2 = lam * v20, m2 = mu * v20, mm2 = T(2) * m2;
    T t[4][6], ss[6];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const T tr = l2 * (G[k][0][0] + G[k][1][1] + G[k][2][2]);
        t[k][0] = mm2 * G[k][0][0] + tr;
        t[k][1] = mm2 * G[k][1][1] + tr;
        t[k][2] = mm2 * G[k][2][2] + tr;
        t[k][3] = m2 * (G[k][1][2] + G[k][2][1]);
        t[k][4] = m2 * (G[k][0][2] + G[k][2][0]);
        t[k][5] = m2 * (G[k][0][1] + G[k][1][0]);
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        ss[s] = (t[0][s] + t[1][s]) + (t[2][s] + t[3][s]);
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k][s] += ss[s];
        ss[s] *= T(5);  // sum_m t_m
    }
    // the symmetric entry (a, b) of (xx, yy, zz, yz, xz, xy)
    auto sym = [](int a, int b) { return a == b ? a : 6 - a - b; };
    // corners: y_k = (4 t_k - sum_m t_m) grad L_k
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T acc = T(0);
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int m = kuhn_rank(TT, b);
                const int s = sym(a, b);
                if (k == m + 1) acc += (T(4) * t[k][s] - ss[s]) * ih[b];
                else if (k == m) acc -= (T(4) * t[k][s] - ss[s]) * ih[b];
            }
            y[3 * k + a] = acc;
        }
    // edges (k, j): y = 4 (t_k grad L_j + t_j grad L_k)
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
                if (j == m + 1) acc += t[k][s] * ih[b];
                else if (j == m) acc -= t[k][s] * ih[b];
                if (k == m + 1) acc += t[j][s] * ih[b];
                else if (k == m) acc -= t[j][s] * ih[b];
            }
            y[3 * (4 + e) + a] = T(4) * acc;
        }
    }
}

}  // namespace fem
