// Multifrontal LDL^T on the device (multifrontal.hpp).
//
// Host: the nested-dissection order (order.rs:19-53), its elimination-tree postorder and the
// fundamental supernodes with their rows (symbolic.rs:41-150); each supernode is a front of the
// multifrontal method (cholesky.rs:44-121): its k pivot columns and the m - k rows below them.
// The supernode tree is cut into levels (a front's level is one more than its highest child's),
// and each level runs as a few launches over all of its fronts:
//  1. zero the fronts (dummy pivots get a unit diagonal, below);
//  2. put A's entries of their pivot columns in place;
//  3. add each child's update matrix (the Schur complement left in the child's front), child slot
//     by child slot, so that no entry takes two additions at once and every sum has a fixed order;
//  4. factor the k pivots: a front of at most 64 rows in one workgroup, its rows in registers (the
//     diagonal-block kernel of the dense factor with the pivot count known only at run time); a
//     taller front of at most 48 pivots by a panel workgroup and one tiled update of the rows past
//     its pivots; a front of more pivots by the dense factor's blocked steps (ldlt_panel: the diagonal block, the block
//     column and the trailing update on the FP64 matrix cores), with its pivots padded to a
//     multiple of 64 by unit pivots that couple to nothing, and its other rows after them, so the
//     trailing blocks end as the update matrix.
// Every front stays in device memory (its L, and the update its parent reads in place), so the
// solve runs through the fronts level by level: forward with the children's update vectors added
// in child order, the diagonal, and backward from the roots. Sums have a fixed order: the results
// are the same from run to run.
#include "multifrontal.hpp"

#include <algorithm>
#include <string>

#include "common.hpp"
#include "dense.hpp"

namespace fem {

namespace {

constexpr int MB = 64;           // fronts of at most 64 rows are factored in one workgroup
constexpr int KMED = 48;         // other fronts of at most 48 pivots: a panel workgroup and tiled updates
constexpr double PIVOT = 1e-10;  // cholesky.rs:22
constexpr uint32_t CHUNK = 4096; // items per block of the segmented launches
constexpr uint32_t KSPLIT = 64; // fronts of at least this many pivots: the split solve
constexpr uint64_t KJOIN = 48;   // a small supernode joins its parent when the two hold at most this many pivots

struct FrontDesc {
    uint64_t off;     // the front: row-major N x N (lower triangle used) at F + off
    uint64_t first;   // its first pivot column (the factor's order)
    uint64_t doff;    // its D values
    uint64_t woff;    // its vector in the solve (N values)
    uint64_t roff;    // its rows (the factor's order)
    uint64_t reloff;  // the parent positions of its rows past the pivots
    uint32_t N, k, m, tpos;  // side, pivots, rows, position of the first row past the pivots
    uint32_t kid0, nkids;    // its children in the children list
    int32_t parent;
    uint32_t pad;
};

// A block of a segmented launch: items start .. start + count - 1 of front `front`.
struct Task {
    uint32_t front, start, count, pad;
};

__device__ __forceinline__ uint32_t pos_of(const FrontDesc& f, uint32_t t) { return t < f.k ? t : f.tpos + (t - f.k); }

// Lane t's value to every lane of the wave (t the same in every lane): two v_readlane.
__device__ __forceinline__ double lane_value(double v, uint32_t t) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)t);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)t);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The diagonal block of the pivots j0 .. j0 + jn - 1 of a front into LDS (row r, column c at
// Ld[r][c]; rows padded to 65), in one coalesced pass of the 256 threads, so that the block's
// solve chain reads LDS instead of waiting on a load at each step.
__device__ __forceinline__ void stage_block(const FrontDesc& f, const double* L, uint32_t j0, uint32_t jn,
                                            double (*Ld)[65]) {
    for (uint32_t e = threadIdx.x; e < jn * 64; e += 256) {
        const uint32_t r = e >> 6, c = e & 63;
        if (c < jn) Ld[r][c] = L[(uint64_t)(j0 + r) * f.N + j0 + c];
    }
}

__global__ void __launch_bounds__(256) k_mf_init(const Task* tasks, const FrontDesc* fronts, double* F) {
    const Task tk = tasks[blockIdx.x];
    const FrontDesc f = fronts[tk.front];
    for (uint32_t i = threadIdx.x; i < tk.count; i += 256) {
        const uint64_t e = (uint64_t)tk.start + i;
        const uint32_t r = (uint32_t)(e / f.N), c = (uint32_t)(e % f.N);
        F[f.off + e] = (r == c && r >= f.k && r < f.tpos) ? 1.0 : 0.0;
    }
}

__global__ void __launch_bounds__(256) k_mf_scatter(uint64_t n, const uint64_t* dst, const uint64_t* src,
                                                     const double* val, double* F) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) F[dst[i]] = val[src[i]];
}

// Adds child `front`'s update (its rows past the pivots, lower triangle) into its parent's front.
__global__ void __launch_bounds__(256) k_mf_extend(const Task* tasks, const FrontDesc* fronts, const uint32_t* rel,
                                                    double* F) {
    const Task tk = tasks[blockIdx.x];
    const FrontDesc c = fronts[tk.front];
    const FrontDesc p = fronts[c.parent];
    const uint32_t mu = c.m - c.k;
    const uint32_t* r = rel + c.reloff;
    for (uint32_t i0 = threadIdx.x; i0 < tk.count; i0 += 256) {
        const uint64_t e = (uint64_t)tk.start + i0;
        const uint32_t i = (uint32_t)(e / mu), j = (uint32_t)(e % mu);
        if (j > i) continue;
        F[p.off + (uint64_t)r[i] * p.N + r[j]] += F[c.off + (uint64_t)(c.tpos + i) * c.N + (c.tpos + j)];
    }
}

__device__ __forceinline__ double recip64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// The k pivots of a front of at most 64 rows (N = 64), right-looking, four lanes per row holding
// the row in registers (lane l: columns l, l + 4, ..., l + 60), one barrier per pivot: the order
// of the dense factor's diagonal-block kernel, stopped after k pivots so that the lower triangle
// past them holds the Schur complement. L (unit, strict lower) and the Schur complement go back
// to the front, D to its D values. A failing pivot (at or below 1e-10 |a_jj|) continues with 1
// and leaves its column + 1 (the factor's order) in *status, the smallest over all fronts.
__global__ void __launch_bounds__(256) k_mf_small(const uint32_t* list, const FrontDesc* fronts, const double* dA,
                                                   double* F, double* Dv, unsigned long long* status) {
    __shared__ double col[2][MB];
    __shared__ double piv[MB];
    const FrontDesc f = fronts[list[blockIdx.x]];
    const int tid = threadIdx.x, row = tid >> 2, lane = tid & 3;
    const int k = (int)f.k;
    double* blk = F + f.off;
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 4 * q;
        a[q] = c <= row ? blk[(size_t)row * MB + c] : 0.0;
    }
    if (tid < MB) piv[tid] = tid < k ? PIVOT * fabs(dA[f.first + tid]) : 0.0;
    int bad = 0;
    // the 64 steps unrolled (every register index known at compile time), each taken only below k
    // (k is the same for the whole workgroup, so the barriers inside are reached by all)
#pragma unroll
    for (int j = 0; j < MB; ++j) {
        if (j < k) {
            const int q = j / 4;
            const bool holds = lane == (j & 3);
            double* cj = col[j & 1];
            if (holds && row >= j) cj[row] = a[q];
            __syncthreads();
            double dj = cj[j];
            const bool ok = dj > piv[j];
            bad = bad == 0 && !ok ? j + 1 : bad;
            dj = ok ? dj : 1.0;
            const double inv = recip64(dj);
            const double lr = row > j ? cj[row] * inv : 0.0;
            const double lq = lane > (j & 3) ? lr : 0.0;
#pragma unroll
            for (int qq = 0; qq < 16; ++qq) {
                if (qq < q) continue;
                a[qq] = fma(qq == q ? -lq : -lr, cj[lane + 4 * qq], a[qq]);
            }
            if (holds) {
                if (row > j) a[q] *= inv;
                else if (row == j) a[q] = dj;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 4 * q;
        if (c <= row) blk[(size_t)row * MB + c] = a[q];
        if (c == row && row < k) Dv[f.doff + row] = a[q];
    }
    if (tid == 0 && bad) atomicMin(status, (unsigned long long)(f.first + bad));
}

// The k pivots of a front of more than 64 rows and at most KMED pivots (row-major N = m, pivots
// first), one workgroup: right-looking over the panel of its k columns only (the rows below take
// the updates of the panel columns), L scaled into the panel, D on its diagonal. The rows past
// the pivots take their Schur update afterwards, in one pass (k_mf_update).
__global__ void __launch_bounds__(256) k_mf_panel(const uint32_t* list, const FrontDesc* fronts, const double* dA,
                                                   double* F, double* Dv, unsigned long long* status) {
    const FrontDesc f = fronts[list[blockIdx.x]];
    double* P = F + f.off;
    const uint32_t N = f.N;
    int bad = 0;
    for (uint32_t j = 0; j < f.k; ++j) {
        double d = P[(uint64_t)j * N + j];
        const bool ok = d > PIVOT * fabs(dA[f.first + j]);
        bad = bad == 0 && !ok ? (int)j + 1 : bad;
        d = ok ? d : 1.0;
        const double inv = recip64(d);
        // rows below the pivot: the panel columns past j, up to the row's own
        for (uint32_t i = j + 1 + threadIdx.x; i < f.m; i += 256) {
            double* row = P + (uint64_t)i * N;
            const double a = row[j] * inv;
            const uint32_t cend = i < f.k - 1 ? i : f.k - 1;
            for (uint32_t c = j + 1; c <= cend; ++c) row[c] = fma(-a, P[(uint64_t)c * N + j], row[c]);
        }
        __syncthreads();
        for (uint32_t i = j + 1 + threadIdx.x; i < f.m; i += 256) P[(uint64_t)i * N + j] *= inv;
        if (threadIdx.x == 0) {
            P[(uint64_t)j * N + j] = d;
            Dv[f.doff + j] = d;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && bad) atomicMin(status, (unsigned long long)(f.first + bad));
}

// The Schur update of a panel front's rows past its pivots, one 64 x 64 tile (rows I, columns C
// of the rows past the pivots, C <= I) per workgroup: S_ic -= sum over j of L_ij d_j L_cj, the L
// rows of the tile staged in LDS, 4 x 4 entries per thread, j in increasing order.
__global__ void __launch_bounds__(256) k_mf_update(const Task* tasks, const FrontDesc* fronts, const double* Dv,
                                                    double* F) {
    __shared__ double Li[KMED][MB + 1];
    __shared__ double Lc[KMED][MB + 1];
    const Task tk = tasks[blockIdx.x];
    const FrontDesc f = fronts[tk.front];
    double* P = F + f.off;
    const uint32_t N = f.N, k = f.k, mu = f.m - f.k;
    const uint32_t I0 = tk.start * MB, C0 = tk.count * MB;  // tile rows and columns (rows past the pivots)
    for (uint32_t e = threadIdx.x; e < k * MB; e += 256) {
        const uint32_t j = e / MB, r = e % MB;
        const uint32_t ri = I0 + r, rc = C0 + r;
        Li[j][r] = ri < mu ? P[(uint64_t)(k + ri) * N + j] : 0.0;
        Lc[j][r] = rc < mu ? P[(uint64_t)(k + rc) * N + j] * Dv[f.doff + j] : 0.0;
    }
    __syncthreads();
    const uint32_t r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (uint32_t j = 0; j < k; ++j) {
        double li[4], lc[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) li[a] = Li[j][r0 + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) lc[b] = Lc[j][c0 + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(li[a], lc[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t i = I0 + r0 + a, c = C0 + c0 + b;
            if (i < mu && c <= i) P[(uint64_t)(k + i) * N + (k + c)] -= acc[a][b];
        }
}

// Forward solve of the fronts of a level: the front's vector takes y at its pivots and 0 below,
// then its children's update vectors in child order; the unit lower L of its pivots runs over it
// in blocks of 64 pivots (the first wave solves a block with lane shuffles, then every thread
// takes rows below it: one 64-term sum per row), and its pivots go back to y (z = L^-1 b at those
// columns). What stays below the pivots is the front's update vector for its parent.
__global__ void __launch_bounds__(256) k_mf_forward(const uint32_t* list, const FrontDesc* fronts, const uint32_t* kids,
                                                     const uint32_t* rel, const double* F, double* w, double* y) {
    __shared__ double Ld[64][65];
    const FrontDesc f = fronts[list[blockIdx.x]];
    double* wf = w + f.woff;
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    for (uint32_t t = tid; t < f.m; t += 256) wf[pos_of(f, t)] = t < f.k ? y[f.first + t] : 0.0;
    __syncthreads();
    for (uint32_t q = 0; q < f.nkids; ++q) {
        const FrontDesc c = fronts[kids[f.kid0 + q]];
        const uint32_t* r = rel + c.reloff;
        const double* wc = w + c.woff + c.tpos;
        for (uint32_t i = tid; i < c.m - c.k; i += 256) wf[r[i]] += wc[i];
        __syncthreads();
    }
    const double* L = F + f.off;
    for (uint32_t j0 = 0; j0 < f.k; j0 += 64) {
        const uint32_t jn = f.k - j0 < 64 ? f.k - j0 : 64;
        stage_block(f, L, j0, jn, Ld);
        __syncthreads();
        if (tid < 64) {
            double wl = lane < jn ? wf[j0 + lane] : 0.0;
            for (uint32_t t = 0; t < jn; ++t) {
                const double v = lane_value(wl, t);
                if (lane > t && lane < jn) wl = fma(-Ld[lane][t], v, wl);
            }
            if (lane < jn) wf[j0 + lane] = wl;
        }
        __syncthreads();
        for (uint32_t t = j0 + jn + tid; t < f.m; t += 256) {
            const uint32_t p = pos_of(f, t);
            const double* lr = L + (uint64_t)p * f.N + j0;
            double sum = 0.0;
            for (uint32_t u = 0; u < jn; ++u) sum = fma(lr[u], wf[j0 + u], sum);
            wf[p] -= sum;
        }
        __syncthreads();
    }
    for (uint32_t t = tid; t < f.k; t += 256) y[f.first + t] = wf[t];
}

// Backward solve of the fronts of a level, in blocks of 64 pivots from the last: each block's
// sums over the rows below it (x_t of the pivots past the block and of the rows past the pivots,
// whose fronts, higher in the tree, are done) run over the four waves in a fixed split, met in a
// fixed order; then the first wave solves the block's L^T with lane shuffles:
// x_j = z_j / d_j - sum over t > j of L_tj x_t.
__global__ void __launch_bounds__(256) k_mf_backward(const uint32_t* list, const FrontDesc* fronts,
                                                      const uint32_t* rows, const double* F, const double* Dv, double* y) {
    __shared__ double red[4][64];
    __shared__ double Ld[64][65];
    const FrontDesc f = fronts[list[blockIdx.x]];
    const double* L = F + f.off;
    const uint32_t* rw = rows + f.roff;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (f.k == 0) return;
    for (int64_t jb = (int64_t)((f.k - 1) / 64) * 64; jb >= 0; jb -= 64) {
        const uint32_t j0 = (uint32_t)jb, jn = f.k - j0 < 64 ? f.k - j0 : 64;
        // sixteen rows in flight per wave (independent partial sums, met in a fixed order)
        double s16[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) s16[q] = 0.0;
        if (lane < jn)
            for (uint32_t t = j0 + jn + wave; t < f.m; t += 64) {  // a ragged last group is predicated
                double xt[16], lv[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const uint32_t tq = t + 4 * q;
                    const bool in = tq < f.m;
                    xt[q] = !in ? 0.0 : tq < f.k ? y[f.first + tq] : y[rw[tq]];
                    lv[q] = in ? L[(uint64_t)pos_of(f, tq) * f.N + j0 + lane] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 16; ++q) s16[q] = fma(lv[q], xt[q], s16[q]);
            }
#pragma unroll
        for (int h = 8; h > 0; h >>= 1)
#pragma unroll
            for (int q = 0; q < h; ++q) s16[q] += s16[q + h];
        red[wave][lane] = s16[0];
        stage_block(f, L, j0, jn, Ld);
        __syncthreads();
        if (tid < 64) {
            const double below = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
            double xv = lane < jn ? y[f.first + j0 + lane] / Dv[f.doff + j0 + lane] - below : 0.0;
            for (int t = (int)jn - 1; t >= 0; --t) {
                const double xt = lane_value(xv, (uint32_t)t);
                if ((int)lane < t) xv = fma(-Ld[t][lane], xt, xv);
            }
            if (lane < jn) y[f.first + j0 + lane] = xv;
        }
        __syncthreads();
    }
}

// The split solve of the fronts with at least KSPLIT pivots: one launch per block of 64 pivots
// over all such fronts of a level (blockIdx.y), the rows below the block in chunks of 64 rows,
// one chunk a workgroup (blockIdx.x), so that the top of the tree, where a level holds a few large
// fronts, streams its factor over many CUs instead of one workgroup a front. The workgroup of the
// chunk right below a block holds the next block's rows, so it also solves that block within the
// same launch: no workgroup waits for another. Sums run in a fixed order (a fixed butterfly in the
// forward steps; the four waves' sums, then the chunks in order, in the backward ones).

// Forward solve of the 64 pivots from j0 of a front's vector wf (its first wave): the unit lower L
// of the block (staged in Ld), z to wf and to y.
__device__ __forceinline__ void mfs_forward_block(const FrontDesc& f, const double (*Ld)[65], double* wf, double* y,
                                                  uint32_t j0, uint32_t lane) {
    const uint32_t jn = f.k - j0 < 64 ? f.k - j0 : 64;
    double wl = lane < jn ? wf[j0 + lane] : 0.0;
    for (uint32_t t = 0; t < jn; ++t) {
        const double v = lane_value(wl, t);
        if (lane > t && lane < jn) wl = fma(-Ld[lane][t], v, wl);
    }
    if (lane < jn) {
        wf[j0 + lane] = wl;
        y[f.first + j0 + lane] = wl;
    }
}

// Start of the forward solve of a split front: its vector (y at its pivots, 0 below, then its
// children's update vectors in child order, as k_mf_forward) and its first block.
__global__ void __launch_bounds__(256) k_mfs_forward_start(const uint32_t* list, const FrontDesc* fronts,
                                                           const uint32_t* kids, const uint32_t* rel, const double* F,
                                                           double* w, double* y) {
    __shared__ double Ld[64][65];
    const FrontDesc f = fronts[list[blockIdx.x]];
    double* wf = w + f.woff;
    const uint32_t tid = threadIdx.x;
    stage_block(f, F + f.off, 0, f.k < 64 ? f.k : 64, Ld);
    for (uint32_t t = tid; t < f.m; t += 256) wf[pos_of(f, t)] = t < f.k ? y[f.first + t] : 0.0;
    __syncthreads();
    for (uint32_t q = 0; q < f.nkids; ++q) {
        const FrontDesc c = fronts[kids[f.kid0 + q]];
        const uint32_t* r = rel + c.reloff;
        const double* wc = w + c.woff + c.tpos;
        for (uint32_t i = tid; i < c.m - c.k; i += 256) wf[r[i]] += wc[i];
        __syncthreads();
    }
    __syncthreads();
    if (tid < 64) mfs_forward_block(f, Ld, wf, y, 0, tid);
}

// Forward step `blk` of the split fronts: the rows of chunk blockIdx.x below the block take
// L[row, block] . z[block] (lane = row; wave w sums the columns 16 w .. 16 w + 15 of the block,
// and the four sums meet in a fixed order); the first chunk then solves the next block.
__global__ void __launch_bounds__(256) k_mfs_forward_step(const uint32_t* list, const FrontDesc* fronts,
                                                          const double* F, double* w, double* y, uint32_t blk) {
    __shared__ double zs[64];
    __shared__ double red[4][64];
    __shared__ double Ld[64][65];
    const FrontDesc f = fronts[list[blockIdx.y]];
    const uint32_t j0 = blk * 64;
    if (j0 >= f.k) return;
    const uint32_t jn = f.k - j0 < 64 ? f.k - j0 : 64;
    const uint32_t r0 = j0 + jn + blockIdx.x * 64;
    if (r0 >= f.m) return;
    const uint32_t rn = f.m - r0 < 64 ? f.m - r0 : 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* L = F + f.off;
    double* wf = w + f.woff;
    if (tid < 64) zs[lane] = lane < jn ? wf[j0 + lane] : 0.0;
    const bool next = blockIdx.x == 0 && j0 + jn < f.k;  // this chunk also solves the next block
    const uint32_t jn1 = f.k - j0 - jn < 64 ? f.k - j0 - jn : 64;
    if (next) stage_block(f, L, j0 + jn, jn1, Ld);
    __syncthreads();
    const uint32_t p = lane < rn ? pos_of(f, r0 + lane) : 0;
    double acc = 0.0;
    if (lane < rn) {
        const double* lr = L + (uint64_t)p * f.N + j0 + wave * 16;
        const uint32_t u1 = jn > wave * 16 ? (jn - wave * 16 < 16 ? jn - wave * 16 : 16) : 0;
        if (u1 == 16) {
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) acc = fma(lr[u], zs[wave * 16 + u], acc);
        } else {
            for (uint32_t u = 0; u < u1; ++u) acc = fma(lr[u], zs[wave * 16 + u], acc);
        }
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (tid < 64 && lane < rn) wf[p] -= (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (next) {
        __syncthreads();
        if (tid < 64) mfs_forward_block(f, Ld, wf, y, j0 + jn, lane);
    }
}

// The number of 64-row chunks below block `blk` of a front.
__device__ __forceinline__ uint32_t mfs_chunks(const FrontDesc& f, uint32_t blk) {
    const uint32_t j0 = blk * 64, jn = f.k - j0 < 64 ? f.k - j0 : 64;
    return (f.m - j0 - jn + 63) / 64;
}

// Backward step `st` of the split fronts (st = 0 takes each front's last block): block b = its
// number of blocks - 1 - st. The first chunk first finishes block b + 1 (the last step's chunk sums
// in chunk order, then L^T of the block: x_j = z_j / d_j - sum over t > j of L_tj x_t); then each
// chunk's sums over its rows, L[t, block b] x_t per column, go to part. The step after a front's
// block 0 only finishes that block.
__global__ void __launch_bounds__(256) k_mfs_backward_step(const uint32_t* list, const FrontDesc* fronts,
                                                           const uint32_t* rows, const double* F, const double* Dv,
                                                           double* y, double* part, uint32_t maxch, uint32_t st) {
    __shared__ double red[4][64];
    __shared__ double Ld[64][65];
    const FrontDesc f = fronts[list[blockIdx.y]];
    const int nb = (int)((f.k + 63) / 64);
    const int b = nb - 1 - (int)st;
    if (b < -1) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* L = F + f.off;
    double* mine = part + (uint64_t)blockIdx.y * 2 * maxch * 64;
    if (blockIdx.x == 0 && b + 1 < nb) {  // finish block b + 1
        const uint32_t j0 = (uint32_t)(b + 1) * 64, jn = f.k - j0 < 64 ? f.k - j0 : 64;
        stage_block(f, L, j0, jn, Ld);
        __syncthreads();
        if (tid < 64) {
            const double* prev = mine + (uint64_t)((st - 1) & 1) * maxch * 64;
            const uint32_t nc = mfs_chunks(f, (uint32_t)(b + 1));
            double below = 0.0;
            uint32_t c = 0;
            for (; c + 8 <= nc; c += 8) {  // eight loads in flight, added in chunk order
                double v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = prev[(c + q) * 64 + lane];
#pragma unroll
                for (int q = 0; q < 8; ++q) below += v[q];
            }
            for (; c < nc; ++c) below += prev[c * 64 + lane];
            double xv = lane < jn ? y[f.first + j0 + lane] / Dv[f.doff + j0 + lane] - below : 0.0;
            for (int t = (int)jn - 1; t >= 0; --t) {
                const double xt = lane_value(xv, (uint32_t)t);
                if ((int)lane < t) xv = fma(-Ld[t][lane], xt, xv);
            }
            if (lane < jn) y[f.first + j0 + lane] = xv;
        }
        __syncthreads();
    }
    if (b < 0) return;
    const uint32_t j0 = (uint32_t)b * 64, jn = f.k - j0 < 64 ? f.k - j0 : 64;
    const uint32_t r0 = j0 + jn + blockIdx.x * 64;
    if (r0 >= f.m) return;
    const uint32_t rn = f.m - r0 < 64 ? f.m - r0 : 64;
    const uint32_t* rw = rows + f.roff;
    double acc = 0.0;
    if (lane < jn) {  // sixteen rows a wave, their loads in flight together (a ragged chunk predicated)
        double xt[16], lv[16];
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint32_t t = r0 + wave + 4 * i;
            const bool in = wave + 4 * i < rn;
            xt[i] = !in ? 0.0 : t < f.k ? y[f.first + t] : y[rw[t]];
            lv[i] = in ? L[(uint64_t)pos_of(f, t) * f.N + j0 + lane] : 0.0;
        }
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) acc = fma(lv[i], xt[i], acc);
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (tid < 64)
        mine[(uint64_t)(st & 1) * maxch * 64 + blockIdx.x * 64 + lane] =
            (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

__global__ void __launch_bounds__(256) k_mf_gather(uint64_t n, const uint64_t* perm, const double* b, double* y) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = b[perm[i]];
}
__global__ void __launch_bounds__(256) k_mf_place(uint64_t n, const uint64_t* perm, const double* y, double* x) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[perm[i]] = y[i];
}

unsigned blocks_of(uint64_t n) { return (unsigned)((n + 255) / 256); }


template <class T>
DArray<T> upload(const std::vector<T>& v, hipStream_t s) {
    DArray<T> d(std::max<size_t>(v.size(), 1));
    if (!v.empty()) d.upload(v.data(), v.size(), s);
    return d;
}

}  // namespace

struct Multifrontal::Impl {
    uint64_t n = 0;
    sym::Counts counts;
    uint64_t ns = 0, nlevels = 0, pool = 0;
    // per level: its fronts (small and large), its tasks and scatter entries
    std::vector<uint64_t> lv_small, lv_med, lv_all;       // offsets into the lists (nlevels + 1)
    std::vector<uint32_t> all_list, small_list, med_list;  // host copies
    // the solve: per level the fronts of one workgroup each, and the split fronts with their
    // largest number of blocks and of chunks below a block
    std::vector<uint64_t> lv_solve, lv_split;
    std::vector<uint32_t> split_blocks, split_chunks;
    DArray<uint32_t> d_solve, d_split;
    DArray<double> part;
    std::vector<std::vector<uint32_t>> lv_large;
    std::vector<FrontDesc> fr;
    DArray<FrontDesc> d_fr;
    DArray<uint32_t> d_all, d_small, d_med, d_kids, d_rel, d_rows;
    DArray<uint64_t> d_perm;
    DArray<double> F, Dv, w, y;

    void build(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, hipStream_t s, uint64_t budget);
};

void Multifrontal::Impl::build(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, hipStream_t s,
                               uint64_t budget) {
    n = a.n;
    const sym::Supernodal S = sym::amalgamate(sym::supernodal(a, sym::nested_dissection(a, xyz)), KJOIN);
    counts = S.counts;
    ns = S.supernodes();
    // fronts: sizes, offsets, levels, children
    fr.resize(ns);
    std::vector<uint32_t> level(ns, 0);
    std::vector<std::vector<uint32_t>> kids(ns);
    for (uint64_t c = 0; c < ns; ++c)
        if (S.sparent[c] >= 0) kids[(size_t)S.sparent[c]].push_back((uint32_t)c);
    uint64_t off = 0, doff = 0, woff = 0, reloff = 0;
    std::vector<uint32_t> kidlist;
    for (uint64_t q = 0; q < ns; ++q) {
        FrontDesc& f = fr[q];
        f.k = (uint32_t)(S.start[q + 1] - S.start[q]);
        f.m = (uint32_t)(S.rptr[q + 1] - S.rptr[q]);
        f.first = S.start[q];
        f.roff = S.rptr[q];
        if (f.m <= (uint32_t)MB) {
            f.N = MB;
            f.tpos = f.k;
        } else if (f.k <= (uint32_t)KMED) {  // a panel front: rows in order, no padding
            f.N = f.m;
            f.tpos = f.k;
        } else {
            const uint32_t kp = (f.k + MB - 1) / MB * MB, mp = (f.m - f.k + MB - 1) / MB * MB;
            f.N = kp + mp;
            f.tpos = kp;
        }
        f.off = off;
        off += (uint64_t)f.N * f.N;
        f.doff = doff;
        doff += f.m <= (uint32_t)MB ? (uint64_t)MB : (uint64_t)f.tpos;  // a panel front: tpos = k
        f.woff = woff;
        woff += f.N;
        f.reloff = reloff;
        reloff += f.m - f.k;
        f.parent = (int32_t)S.sparent[q];
        f.kid0 = (uint32_t)kidlist.size();
        f.nkids = (uint32_t)kids[q].size();
        kidlist.insert(kidlist.end(), kids[q].begin(), kids[q].end());
        for (uint32_t c : kids[q]) level[q] = std::max(level[q], level[c] + 1);
        nlevels = std::max<uint64_t>(nlevels, level[q] + 1);
    }
    pool = off;
    if (pool * 8 > budget)
        throw Error("the multifrontal factor needs " + std::to_string(pool * 8 >> 20) + " MiB of fronts, more than its budget of " +
                    std::to_string(budget >> 20) + " MiB");
    // the parent positions of each front's rows past its pivots
    std::vector<uint32_t> rel(std::max<uint64_t>(reloff, 1));
    for (uint64_t c = 0; c < ns; ++c) {
        const FrontDesc& f = fr[c];
        if (f.parent < 0) continue;
        const FrontDesc& p = fr[(size_t)f.parent];
        const uint64_t* prow = S.rows.data() + S.rptr[(size_t)f.parent];
        for (uint32_t i = 0; i < f.m - f.k; ++i) {
            const uint64_t r = S.rows[S.rptr[c] + f.k + i];
            const uint32_t t = (uint32_t)(std::lower_bound(prow, prow + p.m, r) - prow);
            if (t >= p.m || prow[t] != r) throw Error("internal: a child's row is not in its parent");
            rel[f.reloff + i] = t < p.k ? t : p.tpos + (t - p.k);
        }
    }
    // the fronts of each level
    std::vector<std::vector<uint32_t>> by_level(nlevels);
    for (uint64_t q = 0; q < ns; ++q) by_level[level[q]].push_back((uint32_t)q);
    lv_all.assign(nlevels + 1, 0);
    lv_small.assign(nlevels + 1, 0);
    lv_med.assign(nlevels + 1, 0);
    lv_large.assign(nlevels, {});
    for (uint64_t l = 0; l < nlevels; ++l) {
        for (uint32_t q : by_level[l]) {
            all_list.push_back(q);
            if (fr[q].m <= (uint32_t)MB) small_list.push_back(q);
            else if (fr[q].k <= (uint32_t)KMED) med_list.push_back(q);
            else lv_large[l].push_back(q);
        }
        lv_all[l + 1] = all_list.size();
        lv_small[l + 1] = small_list.size();
        lv_med[l + 1] = med_list.size();
    }
    // the solve's lists
    {
        std::vector<uint32_t> solve_list, split_list;
        lv_solve.assign(nlevels + 1, 0);
        lv_split.assign(nlevels + 1, 0);
        split_blocks.assign(nlevels, 0);
        split_chunks.assign(nlevels, 0);
        uint64_t part_len = 1;
        for (uint64_t l = 0; l < nlevels; ++l) {
            for (uint32_t q : by_level[l]) {
                const FrontDesc& f = fr[q];
                if (f.k >= KSPLIT) {
                    split_list.push_back(q);
                    split_blocks[l] = std::max(split_blocks[l], (f.k + 63) / 64);
                    const uint32_t jn = f.k < 64 ? f.k : 64;
                    split_chunks[l] = std::max(split_chunks[l], (f.m - jn + 63) / 64);
                } else {
                    solve_list.push_back(q);
                }
            }
            lv_solve[l + 1] = solve_list.size();
            lv_split[l + 1] = split_list.size();
            split_chunks[l] = std::max<uint32_t>(split_chunks[l], 1);
            part_len = std::max<uint64_t>(part_len, (lv_split[l + 1] - lv_split[l]) * 2 * split_chunks[l] * 64);
        }
        d_solve = upload(solve_list, s);
        d_split = upload(split_list, s);
        part.resize(part_len);
    }
    // device structures
    d_fr = upload(fr, s);
    d_all = upload(all_list, s);
    d_small = upload(small_list, s);
    d_med = upload(med_list, s);
    d_kids = upload(kidlist, s);
    d_rel = upload(rel, s);
    std::vector<uint32_t> rows32(S.rows.begin(), S.rows.end());
    d_rows = upload(rows32, s);
    d_perm = upload(S.perm, s);
    std::vector<double> dA(n, 0.0);  // A's diagonal in the factor's order (the pivot test)
    for (uint64_t old = 0; old < n; ++old)
        for (uint64_t e = a.ptr[old]; e < a.ptr[old + 1]; ++e)
            if (a.idx[e] == old) dA[S.iperm[old]] = a.val[e];
    DArray<double> d_dA = upload(dA, s);
    DArray<double> d_val = upload(a.val, s);
    F.resize(std::max<uint64_t>(pool, 1));
    Dv.resize(std::max<uint64_t>(doff, 1));
    w.resize(std::max<uint64_t>(woff, 1));
    y.resize(std::max<uint64_t>(n, 1));
    // workspace of the large fronts' blocked steps
    uint32_t maxN = 0, maxkp = 0;
    for (const FrontDesc& f : fr)
        if (f.m > (uint32_t)MB && f.k > (uint32_t)KMED) {
            maxN = std::max(maxN, f.N);
            maxkp = std::max(maxkp, f.tpos);
        }
    DArray<double> P((size_t)std::max<uint32_t>(maxN, 1) * MB), bk((size_t)MB * MB),
        linv((size_t)std::max<uint32_t>(maxkp, 1) * MB), pd(std::max<uint32_t>(maxkp, 1));
    std::vector<uint32_t> large_all;
    for (const auto& l : lv_large) large_all.insert(large_all.end(), l.begin(), l.end());
    DArray<int> lstat(std::max<size_t>(large_all.size(), 1));
    DArray<unsigned long long> sstat(1);
    {
        std::vector<int> z(std::max<size_t>(large_all.size(), 1), 0);
        lstat.upload(z.data(), z.size(), s);
        const unsigned long long none = ~0ull;
        sstat.upload(&none, 1, s);
    }
    // the pivot thresholds' diagonals of each large front (its pivots, then 1 for the unit ones)
    std::vector<double> pdiag;
    std::vector<uint64_t> pdoff(ns, 0);
    for (uint32_t q : large_all) {
        pdoff[q] = pdiag.size();
        for (uint32_t t = 0; t < fr[q].tpos; ++t) pdiag.push_back(t < fr[q].k ? dA[fr[q].first + t] : 1.0);
    }
    DArray<double> d_pdiag = upload(pdiag, s);
    // every level's work lists, built once and uploaded together: the blocks that zero its fronts,
    // A's entries of its pivot columns (cholesky.rs:67-77: the rows at or below each column), and the
    // blocks that add its children's updates, child slot by child slot
    std::vector<Task> init_tasks, ext_tasks;
    std::vector<uint64_t> init_at(nlevels + 1, 0), sc_at(nlevels + 1, 0);
    std::vector<std::vector<uint64_t>> ext_at(nlevels);  // per level: slot s is ext_tasks[ext_at[l][s] .. ext_at[l][s + 1])
    std::vector<uint64_t> dst, src;
    for (uint64_t l = 0; l < nlevels; ++l) {
        const uint32_t* fl = all_list.data() + lv_all[l];
        const uint64_t nf = lv_all[l + 1] - lv_all[l];
        uint32_t maxkids = 0;
        for (uint64_t i = 0; i < nf; ++i) {
            const FrontDesc& f = fr[fl[i]];
            const uint64_t items = (uint64_t)f.N * f.N;
            if (items >= (1ull << 32)) throw Error("a front of the multifrontal factor is too large");
            for (uint64_t st = 0; st < items; st += CHUNK)
                init_tasks.push_back({fl[i], (uint32_t)st, (uint32_t)std::min<uint64_t>(CHUNK, items - st), 0});
            const uint64_t* frow = S.rows.data() + f.roff;
            for (uint32_t jj = 0; jj < f.k; ++jj) {
                const uint64_t j = f.first + jj, old = S.perm[j];
                for (uint64_t e = a.ptr[old]; e < a.ptr[old + 1]; ++e) {
                    const uint64_t r = S.iperm[a.idx[e]];
                    if (r < j) continue;
                    const uint32_t t = (uint32_t)(std::lower_bound(frow, frow + f.m, r) - frow);
                    const uint32_t pr = t < f.k ? t : f.tpos + (t - f.k);
                    dst.push_back(f.off + (uint64_t)pr * f.N + jj);
                    src.push_back(e);
                }
            }
            maxkids = std::max(maxkids, f.nkids);
        }
        init_at[l + 1] = init_tasks.size();
        sc_at[l + 1] = dst.size();
        ext_at[l].push_back(ext_tasks.size());
        for (uint32_t slot = 0; slot < maxkids; ++slot) {
            for (uint64_t i = 0; i < nf; ++i) {
                const FrontDesc& f = fr[fl[i]];
                if (slot >= f.nkids) continue;
                const uint32_t c = kidlist[f.kid0 + slot];
                const uint64_t mu = fr[c].m - fr[c].k, items = mu * mu;
                for (uint64_t st = 0; st < items; st += CHUNK)
                    ext_tasks.push_back({c, (uint32_t)st, (uint32_t)std::min<uint64_t>(CHUNK, items - st), 0});
            }
            ext_at[l].push_back(ext_tasks.size());
        }
    }
    // the tiles of the panel fronts' Schur updates, level by level (lower-triangle tile pairs)
    std::vector<Task> upd_tasks;
    std::vector<uint64_t> upd_at(nlevels + 1, 0);
    for (uint64_t l = 0; l < nlevels; ++l) {
        for (uint64_t i = lv_med[l]; i < lv_med[l + 1]; ++i) {
            const FrontDesc& f = fr[med_list[i]];
            const uint32_t tiles = (f.m - f.k + MB - 1) / MB;
            for (uint32_t I = 0; I < tiles; ++I)
                for (uint32_t C = 0; C <= I; ++C) upd_tasks.push_back({med_list[i], I, C, 0});
        }
        upd_at[l + 1] = upd_tasks.size();
    }
    DArray<Task> d_init = upload(init_tasks, s), d_ext = upload(ext_tasks, s), d_upd = upload(upd_tasks, s);
    DArray<uint64_t> d_dst = upload(dst, s), d_src = upload(src, s);
    size_t large_index = 0;
    std::vector<uint32_t> large_order;
    for (uint64_t l = 0; l < nlevels; ++l) {
        // 1. zero the fronts, 2. A's entries, 3. the children's updates
        const uint64_t ni = init_at[l + 1] - init_at[l];
        if (ni)
            hipLaunchKernelGGL(k_mf_init, dim3((unsigned)ni), dim3(256), 0, s, (const Task*)d_init.data() + init_at[l],
                               (const FrontDesc*)d_fr.data(), F.data());
        const uint64_t nsc = sc_at[l + 1] - sc_at[l];
        if (nsc)
            hipLaunchKernelGGL(k_mf_scatter, dim3(blocks_of(nsc)), dim3(256), 0, s, nsc,
                               (const uint64_t*)d_dst.data() + sc_at[l], (const uint64_t*)d_src.data() + sc_at[l],
                               (const double*)d_val.data(), F.data());
        for (size_t slot = 0; slot + 1 < ext_at[l].size(); ++slot) {
            const uint64_t ne = ext_at[l][slot + 1] - ext_at[l][slot];
            if (ne)
                hipLaunchKernelGGL(k_mf_extend, dim3((unsigned)ne), dim3(256), 0, s,
                                   (const Task*)d_ext.data() + ext_at[l][slot], (const FrontDesc*)d_fr.data(),
                                   (const uint32_t*)d_rel.data(), F.data());
        }
        FEM_HIP(hipGetLastError());
        // 4. the pivots: small fronts in one launch, large ones by the blocked steps
        const uint64_t nsmall = lv_small[l + 1] - lv_small[l];
        if (nsmall)
            hipLaunchKernelGGL(k_mf_small, dim3((unsigned)nsmall), dim3(256), 0, s,
                               (const uint32_t*)d_small.data() + lv_small[l], (const FrontDesc*)d_fr.data(),
                               (const double*)d_dA.data(), F.data(), Dv.data(), sstat.data());
        const uint64_t nmed = lv_med[l + 1] - lv_med[l];
        if (nmed) {
            hipLaunchKernelGGL(k_mf_panel, dim3((unsigned)nmed), dim3(256), 0, s,
                               (const uint32_t*)d_med.data() + lv_med[l], (const FrontDesc*)d_fr.data(),
                               (const double*)d_dA.data(), F.data(), Dv.data(), sstat.data());
            const uint64_t nu = upd_at[l + 1] - upd_at[l];
            if (nu)
                hipLaunchKernelGGL(k_mf_update, dim3((unsigned)nu), dim3(256), 0, s, (const Task*)d_upd.data() + upd_at[l],
                                   (const FrontDesc*)d_fr.data(), (const double*)Dv.data(), F.data());
        }
        FEM_HIP(hipGetLastError());
        for (uint32_t q : lv_large[l]) {
            const FrontDesc& f = fr[q];
            const int nb = (int)(f.N / MB), kb_end = (int)(f.tpos / MB);
            for (int kb = 0; kb < kb_end; ++kb)
                ldlt_panel(F.data() + f.off, (int)f.N, nb, kb, d_pdiag.data() + pdoff[q], Dv.data() + f.doff,
                           linv.data(), bk.data(), P.data(), lstat.data() + large_index, s);
            large_order.push_back(q);
            ++large_index;
        }
    }
    // the first failing pivot in the factor's order (the reference stops at the first, cholesky.rs:92-99)
    unsigned long long first_bad = 0;
    sstat.download(&first_bad, 1, s);
    std::vector<int> lbad(std::max<size_t>(large_all.size(), 1), 0);
    lstat.download(lbad.data(), lbad.size(), s);
    FEM_HIP(hipStreamSynchronize(s));
    uint64_t bad = first_bad == ~0ull ? UINT64_MAX : (uint64_t)first_bad - 1;
    for (size_t i = 0; i < large_order.size(); ++i)
        if (lbad[i]) bad = std::min<uint64_t>(bad, fr[large_order[i]].first + (uint64_t)(lbad[i] - 1));
    if (bad != UINT64_MAX) {
        const uint64_t column = S.perm[bad];
        throw Error("the matrix is not positive definite: the pivot of column " + std::to_string(column) +
                    " is 0 or less to the precision of the arithmetic (its diagonal is " + rust_e(dA[bad]) + ")");
    }
}

Multifrontal::Multifrontal(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, hipStream_t s,
                           uint64_t budget)
    : p_(std::make_unique<Impl>()) {
    p_->build(a, xyz, s, budget);
}
Multifrontal::~Multifrontal() = default;

void Multifrontal::solve(const double* b, double* x, hipStream_t s) const {
    Impl& I = *p_;
    const uint64_t n = I.n;
    if (n == 0) return;
    hipLaunchKernelGGL(k_mf_gather, dim3(blocks_of(n)), dim3(256), 0, s, n, (const uint64_t*)I.d_perm.data(), b,
                       I.y.data());
    const FrontDesc* fr = (const FrontDesc*)I.d_fr.data();
    const double* F = (const double*)I.F.data();
    for (uint64_t l = 0; l < I.nlevels; ++l) {
        const uint64_t nf = I.lv_solve[l + 1] - I.lv_solve[l];
        if (nf)
            hipLaunchKernelGGL(k_mf_forward, dim3((unsigned)nf), dim3(256), 0, s,
                               (const uint32_t*)I.d_solve.data() + I.lv_solve[l], fr, (const uint32_t*)I.d_kids.data(),
                               (const uint32_t*)I.d_rel.data(), F, I.w.data(), I.y.data());
        const uint64_t nsp = I.lv_split[l + 1] - I.lv_split[l];
        if (nsp) {
            const uint32_t* list = (const uint32_t*)I.d_split.data() + I.lv_split[l];
            hipLaunchKernelGGL(k_mfs_forward_start, dim3((unsigned)nsp), dim3(256), 0, s, list, fr,
                               (const uint32_t*)I.d_kids.data(), (const uint32_t*)I.d_rel.data(), F, I.w.data(),
                               I.y.data());
            for (uint32_t blk = 0; blk < I.split_blocks[l]; ++blk)
                hipLaunchKernelGGL(k_mfs_forward_step, dim3(I.split_chunks[l], (unsigned)nsp), dim3(256), 0, s, list, fr,
                                   F, I.w.data(), I.y.data(), blk);
        }
    }
    for (uint64_t l = I.nlevels; l-- > 0;) {
        const uint64_t nsp = I.lv_split[l + 1] - I.lv_split[l];
        if (nsp) {
            const uint32_t* list = (const uint32_t*)I.d_split.data() + I.lv_split[l];
            for (uint32_t st = 0; st <= I.split_blocks[l]; ++st)
                hipLaunchKernelGGL(k_mfs_backward_step, dim3(I.split_chunks[l], (unsigned)nsp), dim3(256), 0, s, list,
                                   fr, (const uint32_t*)I.d_rows.data(), F, (const double*)I.Dv.data(), I.y.data(),
                                   I.part.data(), I.split_chunks[l], st);
        }
        const uint64_t nf = I.lv_solve[l + 1] - I.lv_solve[l];
        if (nf)
            hipLaunchKernelGGL(k_mf_backward, dim3((unsigned)nf), dim3(256), 0, s,
                               (const uint32_t*)I.d_solve.data() + I.lv_solve[l], fr, (const uint32_t*)I.d_rows.data(),
                               F, (const double*)I.Dv.data(), I.y.data());
    }
    hipLaunchKernelGGL(k_mf_place, dim3(blocks_of(n)), dim3(256), 0, s, n, (const uint64_t*)I.d_perm.data(),
                       (const double*)I.y.data(), x);
    FEM_HIP(hipGetLastError());
}

uint64_t Multifrontal::n() const { return p_->n; }
const sym::Counts& Multifrontal::counts() const { return p_->counts; }
uint64_t Multifrontal::front_bytes() const { return p_->pool * 8; }
uint64_t Multifrontal::supernodes() const { return p_->ns; }
uint64_t Multifrontal::levels() const { return p_->nlevels; }

}  // namespace fem
