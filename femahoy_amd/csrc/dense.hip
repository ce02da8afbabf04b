// The dense coarse level on the FP64 matrix cores.
//
// Factor: right-looking blocked LDL^T with 64 x 64 blocks. Each step factors the diagonal block
// in LDS (one workgroup), then forms the block column L_ik = A_ik L_kk^-T D_k^-1 and the trailing
// update A_ij -= (L_ik D_k) L_jk^T as 64 x 64 x 64 products on v_mfma_f64_16x16x4f64.
// Inverse: U = L^-T by a blocked triangular solve on the identity (the same products), then
// A^-1 = (U D^-1/2)(U D^-1/2)^T, one product over the triangular K range of each tile.
// Per cycle the coarse solve is x = A^-1 b, one wave per row, in the V-cycle's precision.
//
// The pivot test is the reference's (crates/fem/src/cholesky.rs:22 and 92-99): a pivot at or
// below 1e-10 of its diagonal in A stops the factor with "the matrix is not positive definite".
#include "dense.hpp"
#include "batch.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

namespace fem {

namespace {

constexpr int NB = 64;
constexpr int KC = 16;
constexpr double PIVOT = 1e-10;

typedef double double4_t __attribute__((ext_vector_type(4)));

struct GemmArgs {
    const double* A;
    int lda;
    const double* B;
    int ldb;
    double* C;
    int ldc;
    int K;
    double alpha, beta;
    int mode;     // 0 rectangle, 1 lower triangle of tiles, 2 symmetric product (K from the tile)
    int ti0, tj0; // tile offsets
    int count;    // mode 1: tiles per side
    double* W;    // optional second output W = C * colscale[col] (ld 64)
    const double* colscale;
};

__global__ void __launch_bounds__(256) k_gemm_nt(GemmArgs g) {
    __shared__ double sA[NB][KC + 1];
    __shared__ double sB[NB][KC + 1];
    int ti, tj;
    if (g.mode == 1) {
        // blockIdx.x -> (a, b) with b <= a
        int t = blockIdx.x, a = 0;
        while (t > a) {
            t -= a + 1;
            ++a;
        }
        ti = g.ti0 + a;
        tj = g.tj0 + t;
    } else {
        ti = g.ti0 + blockIdx.x;
        tj = g.tj0 + blockIdx.y;
    }
    int kofs = 0, K = g.K;
    if (g.mode == 2) {
        kofs = (ti > tj ? ti : tj) * NB;
        K = g.K - kofs;
    }
    const double* Ab = g.A + (size_t)ti * NB * g.lda + kofs;
    const double* Bb = g.B + (size_t)tj * NB * g.ldb + kofs;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    double4_t acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KC) {
#pragma unroll
        for (int e = 0; e < (NB * KC) / 256; ++e) {
            const int idx = tid + 256 * e;
            const int r = idx / KC, c = idx % KC;
            sA[r][c] = Ab[(size_t)r * g.lda + k0 + c];
            sB[r][c] = Bb[(size_t)r * g.ldb + k0 + c];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            double af[2], bf[2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi) af[bi] = sA[wr + bi * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) bf[bj] = sB[wc + bj * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[bi], bf[bj], acc[bi][bj], 0, 0, 0);
        }
        __syncthreads();
    }
    double* Cb = g.C + (size_t)ti * NB * g.ldc + (size_t)tj * NB;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result: lane l, register r holds row (l / 16) + 4 r,
                // column l % 16 of the 16 x 16 block.
                const int row = wr + bi * 16 + (lane >> 4) + 4 * r;
                const int col = wc + bj * 16 + (lane & 15);
                double* c = Cb + (size_t)row * g.ldc + col;
                const double v = (g.alpha != 0.0 ? g.alpha * *c : 0.0) + g.beta * acc[bi][bj][r];
                *c = v;
                if (g.W) g.W[((size_t)ti * NB + row) * NB + col] = v * g.colscale[col];
            }
}

// 1 / x in FP64: the hardware reciprocal refined by two Newton steps (the result is within an
// ulp of the quotient; the IEEE division sequence is about three times as long, and it sits on the
// pivot chain of the factor).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// Factors the diagonal block kb: A_kk = L D L^T (L unit lower). Writes L into the block (strict
// lower part), D into d[kb*64 ..], L^-1 into linv[kb] and D^-1 L^-1 into bk.
// Four lanes per row hold the row in registers (lane l: columns l, l + 4, ..., l + 60). Pivot step
// j publishes the unscaled column j through a double-buffered LDS column behind one barrier (a
// buffer is written again two steps later, after every thread has passed the barrier between), and
// every row r > j takes S[r][c] -= (S[r][j] / d_j) S[c][j] for j < c <= r, then scales S[r][j] by
// 1 / d_j: the order of a right-looking LDL^T. The 64 steps are unrolled so that every register
// index is known at compile time (a register array indexed at run time lives in scratch memory;
// a loop over the steps with selects in its place measured 2.4 times slower), and every update is
// made with few branches: an FMA whose multiplier is 0 where an entry takes no update leaves it
// as it is (a branch per entry cost more than the arithmetic, and a test per entry and step makes
// the compiler hold every step's tests at once, in scratch). The first failing pivot is kept in a
// register and reported once at the end. L^-1 is then formed column by column with four lanes per
// column: lane l sums over the rows k = l mod 4 in four independent partial sums, and the lanes'
// sums meet by two lane exchanges.
__global__ void __launch_bounds__(256) k_ldlt_diag(double* A, int lda, int kb, const double* orig_diag, double* d,
                                                   double* linv, double* bk, int* status) {
    __shared__ double col[2][NB];
    __shared__ double L[NB][NB + 1];
    __shared__ double piv[NB];
    __shared__ double D[NB];
    const int tid = threadIdx.x, row = tid >> 2, lane = tid & 3;
    double* blk = A + (size_t)kb * NB * lda + (size_t)kb * NB;
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 4 * q;
        a[q] = c <= row ? blk[(size_t)row * lda + c] : 0.0;
    }
    if (tid < NB) piv[tid] = PIVOT * fabs(orig_diag[kb * NB + tid]);  // read after the first barrier
    int bad = 0;  // the first failing pivot's column + 1
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int q = j / 4;
        const bool holds = lane == (j & 3);
        double* cj = col[j & 1];
        if (holds && row >= j) cj[row] = a[q];
        __syncthreads();
        double dj = cj[j];
        const bool ok = dj > piv[j];
        bad = bad == 0 && !ok ? kb * NB + j + 1 : bad;
        dj = ok ? dj : 1.0;  // keep going with a harmless pivot; the host reports the failure
        const double inv = recip(dj);
        // rows past the pivot take S[r][c] -= lr S[c][j] on the columns past j: in the group of 4
        // columns holding j the lanes past j's, in the groups above it every lane (the entries
        // above the diagonal, c > r, take it too and are never read); rows up to the pivot take a
        // multiplier of 0, which leaves them as they are
        const double lr = row > j ? cj[row] * inv : 0.0;
        const double lq = lane > (j & 3) ? lr : 0.0;
#pragma unroll
        for (int qq = 0; qq < 16; ++qq) {
            if (qq < q) continue;  // groups of columns below j (known at compile time)
            a[qq] = fma(qq == q ? -lq : -lr, cj[lane + 4 * qq], a[qq]);
        }
        if (holds) {
            if (row > j) a[q] *= inv;
            else if (row == j) a[q] = dj;
        }
    }
    if (tid == 0 && bad && *status == 0) *status = bad;
    // L (unit lower) and D to LDS, the strict lower part of L back to the block
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 4 * q;
        if (c < row) blk[(size_t)row * lda + c] = a[q];
        L[row][c] = c < row ? a[q] : (c == row ? 1.0 : 0.0);
        if (c == row) D[row] = a[q];
    }
    __syncthreads();
    if (tid < NB) d[kb * NB + tid] = D[tid];
    // L^-1: column cc by four lanes, lane l holding rows r = l mod 4 (li[r / 4])
    const int cc = tid >> 2;
    double li[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) li[q] = 0.0;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (4 * q >= r) continue;  // rows k = l + 4 q < r only (known at compile time for 4 q >= r)
            const int k = lane + 4 * q;
            if (k < r) p4[q & 3] += L[r][k] * li[q];
        }
        double part = (p4[0] + p4[1]) + (p4[2] + p4[3]);
        part += __shfl_xor(part, 1, 64);
        part += __shfl_xor(part, 2, 64);
        if (lane == (r & 3)) li[r / 4] = r < cc ? 0.0 : (r == cc ? 1.0 : -part);
    }
    double* lc = linv + (size_t)kb * NB * NB;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = lane + 4 * q;
        lc[r * NB + cc] = li[q];
        bk[r * NB + cc] = li[q] / D[r];
    }
}

__global__ void k_pad(const double* a, int n, double* A, int np) {
    const size_t e = blockIdx.x * (size_t)256 + threadIdx.x;
    if (e >= (size_t)np * np) return;
    const int r = (int)(e / np), c = (int)(e % np);
    A[e] = (r < n && c < n) ? a[(size_t)r * n + c] : (r == c ? 1.0 : 0.0);
}

__global__ void k_diag_of(const double* A, int np, double* dg) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < np) dg[r] = A[(size_t)r * np + r];
}

template <class T>
__global__ void k_zero(T* a, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = T(0);
}

__global__ void k_identity(double* U, int np) {
    const size_t e = blockIdx.x * (size_t)256 + threadIdx.x;
    if (e >= (size_t)np * np) return;
    U[e] = (e / np == e % np) ? 1.0 : 0.0;
}

// V = U D^-1/2 (columns scaled).
__global__ void k_scale_cols(double* U, int np, const double* d) {
    const size_t e = blockIdx.x * (size_t)256 + threadIdx.x;
    if (e >= (size_t)np * np) return;
    U[e] *= 1.0 / sqrt(d[e % np]);
}

// out (n rows, leading dimension ld, zero padded) = the leading n x n block of Ainv.
template <class T>
__global__ void k_extract(const double* Ainv, int np, int n, int ld, T* out) {
    const size_t e = blockIdx.x * (size_t)256 + threadIdx.x;
    if (e >= (size_t)n * ld) return;
    const size_t r = e / ld, c = e % ld;
    out[e] = c < (size_t)n ? (T)Ainv[r * np + c] : T(0);
}

// bc[j] = b[map[j]] for the n dense unknowns (bc is zero beyond n).
template <class T>
__device__ __forceinline__ void coarse_gather_body(unsigned blk, unsigned nblk, int n, const int64_t* __restrict__ map, const T* __restrict__ b, T* __restrict__ bc) {
    const int j = blk * 256 + threadIdx.x;
    if (j < n) bc[j] = b[map[j]];
}
template <class T>
__global__ void k_coarse_gather(int n, const int64_t* __restrict__ map, const T* __restrict__ b, T* __restrict__ bc) {
    coarse_gather_body<T>(blockIdx.x, gridDim.x, n, map, b, bc);
}
template <class T>
struct KCoarseGather {
    static constexpr int NT = 256, MINB = 0;
    int n;
    const int64_t* map;
    const T* b;
    T* bc;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        coarse_gather_body<T>(blk, nblk, n, map, b, bc);
    }
};

// x[map[i]] = sum_j inv[i][j] bc[j]: one wave per row with 16-byte loads, 4 rows per block. Rows
// and bc are zero padded to ld (a multiple of 64), so whole vectors are read.
template <class T>
__device__ __forceinline__ void coarse_gemv_body(unsigned blk, unsigned nblk, const T* __restrict__ inv, int n, int ld, const T* __restrict__ bc,
                                                     const int64_t* __restrict__ map, T* __restrict__ x) {
    constexpr int W = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
    const int row = blk * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const T* a = inv + (size_t)row * ld;
    T s = T(0);
    for (int j = lane * W; j < n; j += 64 * W) {
        T av[W], bv[W];
        if constexpr (W == 4) {
            const float4 A4 = *reinterpret_cast<const float4*>(a + j);
            const float4 B4 = *reinterpret_cast<const float4*>(bc + j);
            av[0] = A4.x; av[1] = A4.y; av[2] = A4.z; av[3] = A4.w;
            bv[0] = B4.x; bv[1] = B4.y; bv[2] = B4.z; bv[3] = B4.w;
        } else {
            const double2 A2 = *reinterpret_cast<const double2*>(a + j);
            const double2 B2 = *reinterpret_cast<const double2*>(bc + j);
            av[0] = A2.x; av[1] = A2.y;
            bv[0] = B2.x; bv[1] = B2.y;
        }
#pragma unroll
        for (int w = 0; w < W; ++w) s += av[w] * bv[w];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[map[row]] = s;
}
template <class T>
__global__ void __launch_bounds__(256) k_coarse_gemv(const T* __restrict__ inv, int n, int ld, const T* __restrict__ bc,
                                                     const int64_t* __restrict__ map, T* __restrict__ x) {
    coarse_gemv_body<T>(blockIdx.x, gridDim.x, inv, n, ld, bc, map, x);
}
template <class T>
struct KCoarseGemv {
    static constexpr int NT = 256, MINB = 0;
    const T* inv;
    int n;
    int ld;
    const T* bc;
    const int64_t* map;
    T* x;
    __device__ __forceinline__ void operator()(unsigned blk, unsigned nblk) const {
        coarse_gemv_body<T>(blk, nblk, inv, n, ld, bc, map, x);
    }
};

// x[i] (= or +=) sum_j inv[i][j] b[j] for i < n: FP64, one wave per row, 4 rows per block. The
// rows of inv have leading dimension ld (a multiple of 64); b has n entries.
template <bool ACC>
__global__ void __launch_bounds__(256) k_dense_gemv(const double* __restrict__ inv, int n, int ld,
                                                    const double* __restrict__ b, double* __restrict__ x) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* a = inv + (size_t)row * ld;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += a[j] * b[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[row] = ACC ? x[row] + s : s;
}

// r = b - A x for the n x n row-major matrix A (leading dimension n): one wave per row.
__global__ void __launch_bounds__(256) k_dense_residual(const double* __restrict__ A, int n, const double* __restrict__ b,
                                                        const double* __restrict__ x, double* __restrict__ r) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* a = A + (size_t)row * n;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += a[j] * x[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) r[row] = b[row] - s;
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

void launch_gemm(const GemmArgs& g, dim3 grid, hipStream_t s) {
    if (grid.x == 0 || grid.y == 0) return;
    hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, s, g);
    check(hipGetLastError(), "k_gemm_nt");
}

}  // namespace

void dense_gemm_nt(int m, int n, int k, double alpha, const double* A, int lda, double beta, const double* B, int ldb,
                   double* C, int ldc, hipStream_t s) {
    if (m % NB || n % NB || k % KC) throw Error("dense_gemm_nt needs sizes in multiples of 64");
    GemmArgs g{A, lda, B, ldb, C, ldc, k, alpha, beta, 0, 0, 0, 0, nullptr, nullptr};
    launch_gemm(g, dim3(m / NB, n / NB), s);
}

void DenseWork::carve(void* base, int n_) {
    n = n_;
    np = (n + NB - 1) / NB * NB;
    nb = np / NB;
    char* p = static_cast<char*>(base);
    auto take = [&](size_t b) {
        char* q = p;
        p += align(b);
        return q;
    };
    A = (double*)take((size_t)np * np * 8);
    U = (double*)take((size_t)np * np * 8);
    inv = (double*)take((size_t)np * np * 8);
    P = (double*)take((size_t)np * NB * 8);
    linv = (double*)take((size_t)nb * NB * NB * 8);
    bk = (double*)take((size_t)NB * NB * 8);
    d = (double*)take((size_t)np * 8);
    dg = (double*)take((size_t)np * 8);
    status = (int*)take(sizeof(int));
}

void ldlt_panel(double* A, int lda, int nb, int kb, const double* dg, double* d, double* linv, double* bk, double* P,
                int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_ldlt_diag, dim3(1), dim3(256), 0, s, A, lda, kb, dg, d, linv, bk, status);
    const int below = nb - kb - 1;
    if (below == 0) return;
    // L_ik = A_ik (D^-1 L^-1)^T ; W_ik = L_ik D_k
    GemmArgs pg{A + (size_t)kb * NB, lda, bk, NB, A + (size_t)kb * NB, lda, NB, 0.0, 1.0, 0,
                kb + 1, 0, 0, P, d + (size_t)kb * NB};
    launch_gemm(pg, dim3(below, 1), s);
    // A_ij -= W_ik L_jk^T for kb < j <= i
    GemmArgs ug{P, NB, A + (size_t)kb * NB, lda, A, lda, NB, 1.0, -1.0, 1, kb + 1, kb + 1, below, nullptr, nullptr};
    launch_gemm(ug, dim3(below * (below + 1) / 2, 1), s);
}

void dense_factor_inverse(const double* a, int n, const DenseWork& w, hipStream_t s) {
    if (n <= 0 || w.n != n) throw Error("internal: the dense workspace is not carved for this size");
    const int nb = w.nb, np = w.np;
    const size_t nn = (size_t)np * np;
    hipLaunchKernelGGL(k_zero<int>, dim3(1), dim3(256), 0, s, w.status, 1);
    const unsigned blocks = (unsigned)((nn + 255) / 256);
    hipLaunchKernelGGL(k_pad, dim3(blocks), dim3(256), 0, s, a, n, w.A, np);
    hipLaunchKernelGGL(k_diag_of, dim3((np + 255) / 256), dim3(256), 0, s, (const double*)w.A, np, w.dg);
    for (int kb = 0; kb < nb; ++kb) ldlt_panel(w.A, np, nb, kb, w.dg, w.d, w.linv, w.bk, w.P, w.status, s);
    int bad = 0;
    FEM_HIP(hipMemcpyAsync(&bad, w.status, sizeof(int), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    if (bad) {
        const int col = bad - 1;
        double diag = 0.0;
        FEM_HIP(hipMemcpyAsync(&diag, w.dg + col, sizeof(double), hipMemcpyDeviceToHost, s));
        FEM_HIP(hipStreamSynchronize(s));
        // cholesky.rs:94-98, the diagonal in Rust's {:e}
        throw Error("the matrix is not positive definite: the pivot of column " + std::to_string(col) +
                    " is 0 or less to the precision of the arithmetic (its diagonal is " + rust_e(diag) + ")");
    }
    // U = L^-T: solve on the identity, working on the transpose (all products are A B^T).
    hipLaunchKernelGGL(k_identity, dim3(blocks), dim3(256), 0, s, w.U, np);
    for (int k = 0; k < nb; ++k) {
        // U_jk = U_jk Linv_kk^T for j <= k
        GemmArgs s1{w.U + (size_t)k * NB, np, w.linv + (size_t)k * NB * NB, NB, w.U + (size_t)k * NB, np,
                    NB, 0.0, 1.0, 0, 0, 0, 0, nullptr, nullptr};
        launch_gemm(s1, dim3(k + 1, 1), s);
        // U_ji -= U_jk L_ik^T for j <= k < i
        const int right = nb - k - 1;
        if (right == 0) continue;
        GemmArgs s2{w.U + (size_t)k * NB, np, w.A + (size_t)k * NB, np, w.U, np, NB, 1.0, -1.0, 0, 0,
                    k + 1, 0, nullptr, nullptr};
        launch_gemm(s2, dim3(k + 1, right), s);
    }
    hipLaunchKernelGGL(k_scale_cols, dim3(blocks), dim3(256), 0, s, w.U, np, (const double*)w.d);
    // A^-1 = V V^T with V upper triangular: K from max(i, j) of the tile.
    GemmArgs fg{w.U, np, w.U, np, w.inv, np, np, 0.0, 1.0, 2, 0, 0, 0, nullptr, nullptr};
    launch_gemm(fg, dim3(nb, nb), s);
    FEM_HIP(hipGetLastError());
}

void dense_solve_refined(const double* a, int n, const DenseWork& w, const double* b, double* x, double* r,
                         hipStream_t s) {
    const unsigned rows = (unsigned)((n + 3) / 4);
    // x = inv b (rows of inv are padded to np with zeros past n: only the first n columns are read)
    hipLaunchKernelGGL((k_dense_gemv<false>), dim3(rows), dim3(256), 0, s, (const double*)w.inv, n, w.np, b, x);
    // r = b - A x on the original matrix, e = inv r, x += e
    hipLaunchKernelGGL(k_dense_residual, dim3(rows), dim3(256), 0, s, a, n, b, (const double*)x, r);
    hipLaunchKernelGGL((k_dense_gemv<true>), dim3(rows), dim3(256), 0, s, (const double*)w.inv, n, w.np, (const double*)r, x);
    FEM_HIP(hipGetLastError());
}

void DenseCoarse::factor(double* a, int n, bool fp32, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    n_ = n;
    fp32_ = fp32;
    Buffer work(DenseWork::bytes(n));
    DenseWork w;
    w.carve(work.get(), n);
    dense_factor_inverse(a, n, w, s);
    np_ = w.np;
    ld_ = w.np;  // rows zero padded to a multiple of 64
    const unsigned eb = (unsigned)(((size_t)n * ld_ + 255) / 256);
    const unsigned zb = (unsigned)((ld_ + 255) / 256);
    if (fp32) {
        inv32_.resize((size_t)n * ld_);
        bc32_.resize(ld_);
        hipLaunchKernelGGL(k_extract<float>, dim3(eb), dim3(256), 0, s, (const double*)w.inv, w.np, n, ld_, inv32_.data());
        hipLaunchKernelGGL(k_zero<float>, dim3(zb), dim3(256), 0, s, bc32_.data(), ld_);
    } else {
        inv64_.resize((size_t)n * ld_);
        bc64_.resize(ld_);
        hipLaunchKernelGGL(k_extract<double>, dim3(eb), dim3(256), 0, s, (const double*)w.inv, w.np, n, ld_, inv64_.data());
        hipLaunchKernelGGL(k_zero<double>, dim3(zb), dim3(256), 0, s, bc64_.data(), ld_);
    }
    FEM_HIP(hipGetLastError());
    FEM_HIP(hipStreamSynchronize(s));  // the workspace is freed at the end of this scope
    factor_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double dn = (double)np_;
    factor_flops_ = dn * dn * dn / 3.0 + dn * dn * dn / 3.0 + dn * dn * dn / 3.0;
}

void DenseCoarse::shadow_of(const DenseCoarse& o, hipStream_t s) {
    n_ = o.n_;
    np_ = o.np_;
    ld_ = o.ld_;
    fp32_ = o.fp32_;
    factor_seconds_ = 0.0;
    factor_flops_ = 0.0;
    inv64_.view(o.inv64_);
    inv32_.view(o.inv32_);
    map_.view(o.map_);
    dense_of_.view(o.dense_of_);
    bc64_.like(o.bc64_, s);
    bc32_.like(o.bc32_, s);
}

template <class T>
void DenseCoarse::solve(const T* b, T* x, hipStream_t s) const {
    const T* inv;
    T* bc;
    if constexpr (sizeof(T) == 4) {
        if (!fp32_) throw Error("the coarse inverse is FP64");
        inv = inv32_.data();
        bc = bc32_.data();
    } else {
        if (fp32_) throw Error("the coarse inverse is FP32");
        inv = inv64_.data();
        bc = bc64_.data();
    }
    run_k(KCoarseGather<T>{n_, map_.data(), b, bc}, (unsigned)((n_ + 255) / 256), s);
    run_k(KCoarseGemv<T>{inv, n_, ld_, (const T*)bc, map_.data(), x}, (unsigned)((n_ + 3) / 4), s);
}

template <class T>
T* DenseCoarse::rhs() const {
    if constexpr (sizeof(T) == 4) {
        if (!fp32_) throw Error("the coarse inverse is FP64");
        return bc32_.data();
    } else {
        if (fp32_) throw Error("the coarse inverse is FP32");
        return bc64_.data();
    }
}

template <class T>
void DenseCoarse::solve_rhs(T* x, hipStream_t s) const {
    const T* inv;
    if constexpr (sizeof(T) == 4) inv = inv32_.data();
    else inv = inv64_.data();
    run_k(KCoarseGemv<T>{inv, n_, ld_, (const T*)rhs<T>(), map_.data(), x}, (unsigned)((n_ + 3) / 4), s);
}

template void DenseCoarse::solve<float>(const float*, float*, hipStream_t) const;
template void DenseCoarse::solve<double>(const double*, double*, hipStream_t) const;
template float* DenseCoarse::rhs<float>() const;
template double* DenseCoarse::rhs<double>() const;
template void DenseCoarse::solve_rhs<float>(float*, hipStream_t) const;
template void DenseCoarse::solve_rhs<double>(double*, hipStream_t) const;

}  // namespace fem

// ---------------------------------------------------------------------------------------------
// Test hooks (include/fem_testing.h).
#include <cstring>

#include "../../include/fem_testing.h"

namespace {
template <class F>
int hook(char* err, size_t errlen, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
}
}  // namespace

extern "C" int fem_test_gemm_nt(int m, int n, int k, const double* A, const double* B, double* C, char* err,
                                size_t errlen) {
    return hook(err, errlen, [&] {
        fem::DArray<double> a((size_t)m * k), b((size_t)n * k), c((size_t)m * n);
        a.upload(A, (size_t)m * k);
        b.upload(B, (size_t)n * k);
        fem::dense_gemm_nt(m, n, k, 0.0, a.data(), k, 1.0, b.data(), k, c.data(), n, nullptr);
        c.download(C, (size_t)m * n);
        FEM_HIP(hipDeviceSynchronize());
    });
}

extern "C" int fem_test_dense_solve32(int n, const double* A, const float* b, float* x, char* err, size_t errlen) {
    return hook(err, errlen, [&] {
        fem::DArray<double> a((size_t)n * n);
        a.upload(A, (size_t)n * n);
        fem::DenseCoarse dc;
        dc.factor(a.data(), n, true, nullptr);
        std::vector<int64_t> id(n);
        for (int i = 0; i < n; ++i) id[i] = i;
        dc.map().resize(n);
        dc.map().upload(id.data(), n);
        fem::DArray<float> bd(n), xd(n);
        bd.upload(b, n);
        dc.solve<float>(bd.data(), xd.data(), nullptr);
        xd.download(x, n);
        FEM_HIP(hipDeviceSynchronize());
    });
}

extern "C" int fem_test_dense_inverse(int n, const double* A, double* Ainv, char* err, size_t errlen) {
    return hook(err, errlen, [&] {
        fem::DArray<double> a((size_t)n * n);
        a.upload(A, (size_t)n * n);
        fem::DenseCoarse dc;
        dc.factor(a.data(), n, false, nullptr);
        // x = inv b for the unit vectors, through the solve path: map = identity
        std::vector<int64_t> id(n);
        for (int i = 0; i < n; ++i) id[i] = i;
        dc.map().resize(n);
        dc.map().upload(id.data(), n);
        fem::DArray<double> e(n), x(n);
        std::vector<double> col(n);
        for (int j = 0; j < n; ++j) {
            std::vector<double> unit(n, 0.0);
            unit[j] = 1.0;
            e.upload(unit.data(), n);
            dc.solve<double>(e.data(), x.data(), nullptr);
            x.download(col.data(), n);
            FEM_HIP(hipDeviceSynchronize());
            for (int i = 0; i < n; ++i) Ainv[(size_t)i * n + j] = col[i];
        }
    });
}
