// Column-marching kernels of the hex8 operator with a scalar coefficient per cell (heat.rs:1246
// conduction matrices, node-centric). One thread owns a column (i, j) of nodes and walks a chunk of
// march_kz<T> planes along z, keeping the 3 x 3 x 3 window of x and the 2 x 2 x 2 cells of k in
// registers, so each step loads one new plane instead of 27 x values and 8 cells per node.
//
// Two load shapes (MarchShape picks one per arithmetic type, from timings on a 129^3 grid):
//  - shuffle: a wave owns 62 consecutive columns plus a halo lane on each side; each lane loads
//    the 3 x values (rows j-1, j, j+1) and 2 cells of its own column and takes the x-neighbours
//    from the lanes beside it with DPP wave shifts (wave_shr:1, wave_shl:1; ds_bpermute before,
//    which made the FP64 product 1-5 % slower). This cuts the vector loads of a plane from 13 to
//    5 per lane. At 129^3 the HBM traffic of a Chebyshev step is 1.04x its algorithmic bytes.
//  - plain: every thread loads the 9 x values and 4 cells of its new plane itself.
#pragma once

#include "kernels.hpp"

namespace fem {

// kz: the planes a thread walks; pf: how many planes ahead of the one computed its loads are
// issued. At 129^3 every wave of the launch is resident at once. Measured at 129^3 (FP32
// Chebyshev step, FP64 q = A p): kz 6, pf 1: 12.4 and 18.3 us; pf 2: 12.4 and 17.9; pf 3: 12.8 and
// 18.5; kz 8, pf 2: 13.2 and 18.1; kz 4, pf 2: 13.5 and 19.1. Non-temporal node loads or stores
// cost 2.5-3.5 us (the level's vectors live in the last-level cache between launches), and
// capping the FP64 kernel at 80 VGPRs for 6 waves per SIMD makes it spill (41 us).
template <class T>
struct MarchShape;
template <>
struct MarchShape<float> {
    static constexpr bool shuffle = true;
    static constexpr int kz = 6;
    static constexpr int pf = 1;
};
template <>
struct MarchShape<double> {
    static constexpr bool shuffle = true;
    static constexpr int kz = 6;
    static constexpr int pf = 1;
};

// Columns a block of BLOCK threads owns.
template <bool SHFL>
constexpr int march_cols_per_block() {
    return SHFL ? (BLOCK / 64) * 62 : BLOCK;
}

// MARCH_CHEB_RZ: a Chebyshev step that also sums r.z over its output z (the preconditioned residual
// of the outer PCG, r in FP64) into block partials.
enum MarchMode { MARCH_APPLY = 0, MARCH_RESIDUAL = 1, MARCH_CHEB = 2, MARCH_PCG_Q = 3, MARCH_PCG_START = 4, MARCH_CHEB_RZ = 5 };

template <class T, class S = T>
struct MarchArgs {
    GridDims g;
    const S* k;     // per cell, stored in S (T, or BF16)
    T c[8];         // conduction pattern coefficients for k = 1
    const T* x;     // input field (p for MARCH_PCG_Q)
    T* y;           // output: y, r, x_out, q
    const T* b;     // right side (residual, cheb, pcg_start)
    T* d;           // Chebyshev direction
    const S* invd;  // D^-1, stored in S
    T c1, c2;       // Chebyshev coefficients
    double* part;   // block partials (pcg_q: p.q; pcg_start: b.b; cheb_rz: r.z)
    double* part2;  // pcg_start: r.r
    const double* r;  // cheb_rz: the outer residual
};

// The value of v held by the lane below (CTRL 0x138, wave_shr:1) or above (0x130, wave_shl:1)
// in the wave, by a DPP move; a lane without that neighbour gets its own value.
template <class T, int CTRL>
__device__ __forceinline__ T lane_shift(T v) {
    if constexpr (sizeof(T) == 4) {
        const int b = __float_as_int(v);
        return __int_as_float(__builtin_amdgcn_update_dpp(b, b, CTRL, 0xf, 0xf, false));
    } else {
        const long long bits = __double_as_longlong(v);
        const int lo = (int)(bits & 0xffffffffll), hi = (int)(bits >> 32);
        const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
        const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
        return __longlong_as_double(((long long)h2 << 32) | (long long)(unsigned)l2);
    }
}

// A raw buffer resource over n values of T (gfx9 descriptor word 3; n * sizeof(T) < 2^31).
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const T* p, size_t n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * sizeof(T)), 0x00020000);
}

// The value at element `idx` of a buffer, or 0 when `valid` is false (the offset then lies past
// the buffer's end and the load returns 0).
template <class T>
__device__ __forceinline__ T buffer_value(__amdgpu_buffer_rsrc_t r, bool valid, unsigned idx) {
    const int off = valid ? (int)(idx * (unsigned)sizeof(T)) : (int)0x80000000u;
    if constexpr (sizeof(T) == 4) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        return __longlong_as_double((long long)(((unsigned long long)v[1] << 32) | v[0]));
    }
}

// The value at element idx of a buffer stored in S, as T, or 0 when valid is false.
template <class T, class S>
__device__ __forceinline__ T buffer_value_as(__amdgpu_buffer_rsrc_t r, bool valid, unsigned idx) {
    if constexpr (sizeof(S) == 2) {
        const int off = valid ? (int)(idx * 2u) : (int)0x80000000u;
        return T(bf16_value((bf16_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
    } else {
        return T(buffer_value<S>(r, valid, idx));
    }
}

// The largest grid (nodes) the marching kernels take: byte offsets of 31 bits.
constexpr size_t MARCH_MAX_BYTES = size_t(1) << 31;

// The tile of a block, XCD-aware: the hardware hands workgroups to the 8 XCDs round-robin by
// linear id, and each XCD has its own L2, so linear id p takes tile (p % 8) * (G / 8) + p / 8.
// Tiles that share halo rows and planes (neighbours in the x-fastest tile order) then run on one
// XCD and meet in its L2. The tail beyond a multiple of 8 keeps its own id.
__device__ __forceinline__ void xcd_tile(unsigned& bx, unsigned& by) {
    const unsigned G = gridDim.x * gridDim.y;
    const unsigned p = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned per = G / 8;
    const unsigned t = p < per * 8 ? (p % 8) * per + p / 8 : p;
    bx = t % gridDim.x;
    by = t / gridDim.x;
}

template <class T, int MODE, bool XCD = true, bool SHFL = MarchShape<T>::shuffle, int KZ = MarchShape<T>::kz, class S = T,
          int PF = MarchShape<T>::pf>
__global__ void __launch_bounds__(BLOCK) k_march_hex8(MarchArgs<T, S> a) {
    const GridDims g = a.g;
    unsigned bx = blockIdx.x, by = blockIdx.y;
    if (XCD) xcd_tile(bx, by);
    const long long cols = (long long)g.nx * g.ny;
    const int lane = threadIdx.x & 63;
    long long col;
    bool own;
    if constexpr (SHFL) {
        const long long wave = bx * (long long)(BLOCK / 64) + (threadIdx.x >> 6);
        col = wave * 62 + lane - 1;
        own = lane >= 1 && lane <= 62 && col < cols;
    } else {
        col = bx * (long long)BLOCK + threadIdx.x;
        own = col < cols;
    }
    const bool incol = col >= 0 && col < cols;
    const long long cc = incol ? col : 0;
    const int i = (int)(cc % g.nx), j = (int)(cc / g.nx);
    const int k0 = by * KZ;
    const int k1 = min(k0 + KZ, g.nz);
    const size_t plane = (size_t)cols;
    const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
    const size_t cplane = (size_t)ncx * ncy;
    // clamped in-plane indices and their validity
    int jj[3];
    bool vj[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        vj[t] = incol && j + t - 1 >= 0 && j + t - 1 < g.ny;
        jj[t] = min(max(j + t - 1, 0), g.ny - 1);
    }
    int cj_[2];
    bool cj[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        cj[t] = incol && j - 1 + t >= 0 && j - 1 + t < ncy;
        cj_[t] = min(max(j - 1 + t, 0), max(ncy - 1, 0));
    }
    const bool hasl = i >= 1, hasr = i + 1 < g.nx;
    // Loads through buffer resources: an out-of-grid neighbour gets an offset past the end of
    // the buffer, which the hardware range check turns into 0 without touching memory, so a plane
    // is loaded without branches or selects.
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(a.x, (size_t)g.nx * g.ny * g.nz);
    const __amdgpu_buffer_rsrc_t rk = make_rsrc(a.k, (size_t)ncx * ncy * ncz);
    // A plane of x and a plane of cells are loaded raw (the lane's own column, or all 9 / 4 values
    // without shuffles) and expanded into the 3 x 3 and 2 x 2 windows when used, so that a load
    // can be issued one plane ahead without waiting on the lane shuffles.
    constexpr int RX = SHFL ? 3 : 9, RC = SHFL ? 2 : 4;
    auto load_plane = [&](int kk, T raw[RX]) {
        const bool vk = kk >= 0 && kk < g.nz;
        const unsigned base = (unsigned)kk * (unsigned)plane;
        if constexpr (SHFL) {
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
                raw[dy] = buffer_value<T>(rx, vk && vj[dy], base + (unsigned)jj[dy] * g.nx + i);
        } else {
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const int p = i + dx - 1;
                    raw[dx + 3 * dy] = buffer_value<T>(rx, vk && vj[dy] && p >= 0 && p < g.nx,
                                                       base + (unsigned)jj[dy] * g.nx + p);
                }
        }
    };
    auto expand_plane = [&](const T raw[RX], T out[9]) {
        if constexpr (SHFL) {
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const T l = lane_shift<T, 0x138>(raw[dy]), r = lane_shift<T, 0x130>(raw[dy]);
                out[0 + 3 * dy] = hasl ? l : T(0);
                out[1 + 3 * dy] = raw[dy];
                out[2 + 3 * dy] = hasr ? r : T(0);
            }
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) out[t] = raw[t];
        }
    };
    auto load_cells = [&](int kc, T raw[RC]) {
        const bool vk = kc >= 0 && kc < ncz;
        const unsigned base = (unsigned)kc * (unsigned)cplane;
        if constexpr (SHFL) {
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
                raw[cy] = buffer_value_as<T, S>(rk, vk && cj[cy] && i < ncx, base + (unsigned)cj_[cy] * ncx + i);
        } else {
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const int p = i - 1 + cx;
                    raw[cx + 2 * cy] = buffer_value_as<T, S>(rk, vk && cj[cy] && p >= 0 && p < ncx,
                                                             base + (unsigned)cj_[cy] * ncx + p);
                }
        }
    };
    auto expand_cells = [&](const T raw[RC], T out[4]) {
        if constexpr (SHFL) {
#pragma unroll
            for (int cy = 0; cy < 2; ++cy) {
                const T l = lane_shift<T, 0x138>(raw[cy]);
                out[0 + 2 * cy] = hasl ? l : T(0);
                out[1 + 2 * cy] = raw[cy];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) out[t] = raw[t];
        }
    };
    // The node data of the update (b; d and D^-1 for a Chebyshev step) of the column at plane kk.
    constexpr bool CHEB = MODE == MARCH_CHEB || MODE == MARCH_CHEB_RZ;
    constexpr int NV = CHEB ? 3 : (MODE == MARCH_RESIDUAL || MODE == MARCH_PCG_START) ? 1 : 0;
    const size_t nodes = (size_t)g.nx * g.ny * g.nz;
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(a.b, nodes), rd = make_rsrc(a.d, nodes), ri = make_rsrc(a.invd, nodes);
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(a.r, nodes);
    auto load_node = [&](int kk, T out[3], double& rv) {
        const bool v = own && kk < k1;
        const unsigned u = (unsigned)kk * (unsigned)plane + (unsigned)col;
        out[0] = NV > 0 ? buffer_value<T>(rb, v, u) : T(0);
        out[1] = NV > 1 ? buffer_value<T>(rd, v, u) : T(0);
        out[2] = NV > 2 ? buffer_value_as<T, S>(ri, v, u) : T(0);
        rv = MODE == MARCH_CHEB_RZ ? buffer_value<double>(rr, v, u) : 0.0;
    };
    // Raw loads by plane: x planes k0 - 1 .. k0 + KZ, cell planes k0 - 1 .. k0 + KZ - 1, node data
    // of planes k0 .. k0 + KZ - 1. Fully unrolled, so each entry is a register; the loads of step
    // st + PF are issued at step st, which keeps PF planes in flight behind the one computed.
    T xr[KZ + 2][RX], kr[KZ + 1][RC], nr[KZ][3];
    double rn[KZ];
    T xv[27], kv[8];  // x planes k-1, k, k+1 (9 each); cells of planes k-1, k (4 each)
    load_plane(k0 - 1, xr[0]);
    load_plane(k0, xr[1]);
    load_cells(k0 - 1, kr[0]);
#pragma unroll
    for (int st = 0; st < PF && st < KZ; ++st) {
        load_plane(k0 + st + 1, xr[st + 2]);
        load_cells(k0 + st, kr[st + 1]);
        load_node(k0 + st, nr[st], rn[st]);
    }
    expand_plane(xr[0], xv);
    expand_plane(xr[1], xv + 9);
    expand_cells(kr[0], kv);
    const bool fxi = ((g.fix & 1u) && i == 0) || ((g.fix & 2u) && i == g.nx - 1);
    const bool fxj = ((g.fix & 4u) && j == 0) || ((g.fix & 8u) && j == g.ny - 1);
    double acc1 = 0.0, acc2 = 0.0;
#pragma unroll
    for (int st = 0; st < KZ; ++st) {
        const int kk = k0 + st;
        if (kk >= k1) break;  // uniform over the block
        if (st + PF < KZ) {
            load_plane(kk + PF + 1, xr[st + PF + 2]);
            load_cells(kk + PF, kr[st + PF + 1]);
            load_node(kk + PF, nr[st + PF], rn[st + PF]);
        }
        expand_plane(xr[st + 2], xv + 18);
        expand_cells(kr[st + 1], kv + 4);
        const T* nv = nr[st];
        T acc = T(0);
#pragma unroll
        for (int cz = 0; cz < 2; ++cz)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cx = 0; cx < 2; ++cx) {
                    const int lx = 1 - cx, ly = 1 - cy, lz = 1 - cz;
                    T s = T(0);
#pragma unroll
                    for (int qz = 0; qz < 2; ++qz)
#pragma unroll
                        for (int qy = 0; qy < 2; ++qy)
#pragma unroll
                            for (int qx = 0; qx < 2; ++qx)
                                s += a.c[(qx ^ lx) | ((qy ^ ly) << 1) | ((qz ^ lz) << 2)] *
                                     xv[(cx + qx) + 3 * (cy + qy) + 9 * (cz + qz)];
                    acc += kv[cx + 2 * cy + 4 * cz] * s;
                }
        if (own) {
            const bool fixed = fxi || fxj || ((g.fix & 16u) && kk == 0) || ((g.fix & 32u) && kk == g.nz - 1);
            const size_t u = (size_t)kk * plane + (size_t)col;
            const T xc = xv[13];
            if (MODE == MARCH_APPLY) {
                a.y[u] = fixed ? T(0) : acc;
            } else if (MODE == MARCH_RESIDUAL) {
                a.y[u] = fixed ? T(0) : nv[0] - acc;
            } else if (CHEB) {
                const T dn = fixed ? T(0) : a.c1 * nv[1] + a.c2 * nv[2] * (nv[0] - acc);
                const T z = fixed ? T(0) : xc + dn;
                a.d[u] = dn;
                a.y[u] = z;
                if (MODE == MARCH_CHEB_RZ) acc1 += rn[st] * (double)z;
            } else if (MODE == MARCH_PCG_Q) {
                const T q = fixed ? T(0) : acc;
                a.y[u] = q;
                acc1 += (double)xc * (double)q;
            } else if (MODE == MARCH_PCG_START) {
                const T bu = nv[0];
                const T r = fixed ? T(0) : bu - acc;
                a.y[u] = r;
                acc1 += (double)bu * (double)bu;
                acc2 += (double)r * (double)r;
            }
        }
#pragma unroll
        for (int t = 0; t < 18; ++t) xv[t] = xv[t + 9];
#pragma unroll
        for (int t = 0; t < 4; ++t) kv[t] = kv[t + 4];
    }
    if (MODE == MARCH_PCG_Q || MODE == MARCH_PCG_START || MODE == MARCH_CHEB_RZ) {
        acc1 = block_sum(acc1);
        if (MODE == MARCH_PCG_START) acc2 = block_sum(acc2);
        if (threadIdx.x == 0) {
            const size_t blk = by * (size_t)gridDim.x + bx;
            a.part[blk] = acc1;
            if (MODE == MARCH_PCG_START) a.part2[blk] = acc2;
        }
    }
}

// The launch grid of the column-marching kernels.
template <class T, bool SHFL = MarchShape<T>::shuffle, int KZ = MarchShape<T>::kz>
inline dim3 march_grid(const GridDims& g) {
    const size_t cols = (size_t)g.nx * g.ny;
    const size_t per = march_cols_per_block<SHFL>();
    return dim3((unsigned)((cols + per - 1) / per), (unsigned)((g.nz + KZ - 1) / KZ));
}

// Blocks of a march launch, the partials a reducing march kernel writes.
template <class T>
inline size_t march_blocks(const GridDims& g) {
    const dim3 d = march_grid<T>(g);
    return (size_t)d.x * d.y;
}

}  // namespace fem
