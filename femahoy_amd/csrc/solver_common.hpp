// What the box solver and the assembled (CSR) solver share on the host: kernel launch helpers,
// the solver options and the reference's number format in error texts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <string>
#include <utility>

#include "../../include/fem_solver.h"
#include "batch.hpp"
#include "kernels.hpp"

namespace fem {

inline unsigned nblocks(size_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// Blocks of a grid-stride reduction over n items (a few items per thread; fewer partials).
inline unsigned rblocks(size_t n) { return nblocks(n) < 2048u ? nblocks(n) : 2048u; }

// Launches `kernel` on `blocks` blocks of BLOCK threads.
// (These plain launches cannot be replayed from a Recorder's record: under one they launch
// nothing and mark the recording unbatchable.)
template <class... KArgs, class... Args>
void launch_blocks(void (*kernel)(KArgs...), unsigned blocks, hipStream_t s, Args&&... args) {
    if (blocks == 0 || recording_unbatchable()) return;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(BLOCK), 0, s, static_cast<KArgs>(args)...);
}

// Launches a partial-sum kernel: one block of RBLOCK threads.
template <class... KArgs, class... Args>
void launch_reduce(void (*kernel)(KArgs...), hipStream_t s, Args&&... args) {
    if (recording_unbatchable()) return;
    hipLaunchKernelGGL(kernel, dim3(1), dim3(RBLOCK), 0, s, static_cast<KArgs>(args)...);
}

// Launches `kernel` with one thread per item of n items.
template <class... KArgs, class... Args>
void launch(void (*kernel)(KArgs...), size_t n, hipStream_t s, Args&&... args) {
    if (n == 0) return;
    launch_blocks(kernel, nblocks(n), s, std::forward<Args>(args)...);
}

// The options of a solver (fem_options after defaults).
struct Options {
    double tol = 1e-10;
    uint64_t max_iter = 500;
    int method = FEM_METHOD_AUTO;
    int precision = FEM_PREC_FP32;
    uint64_t crossover = 3000;
    int levels = 0;
    int degree = 0;  // 0: level_degree's automatic choice
    double cheb_lower = 0.1, cheb_upper = 1.1;
    int power_steps = 20;
    uint64_t coarse_max = 6000;
    int refine = 1;
    int device = 0;
    int coarse_galerkin = 1;
    int crossover_kind = FEM_CROSSOVER_UNKNOWNS;  // solve.rs:28-39
    double crossover_flops = 1.5e5;               // elastic.rs:263 DIRECT_FLOPS
    uint64_t crossover_entries = 250000000;       // elastic.rs:269 DIRECT_ENTRIES
    bool crossover_flat = false;
    int direct_solver = FEM_DIRECT_AUTO;
};

// The smoothing steps (Chebyshev degree, pre and post) of level l of a box hierarchy. An explicit
// degree applies on every level. The automatic choice (degree 0) is 3 on the finest level and 1
// below it for scalar problems, 2 on every level for vector ones: measured fastest on the heat 128^3 box (1015 against
// 945 MDOF/s with 2 everywhere: 8 iterations instead of 9 with cheaper coarse cycles) and on the
// tet10 elasticity box (315 against 236 MDOF/s with 3 and 1, whose coarse tet4 levels need the
// smoothing).
inline int level_degree(int degree, size_t l, bool vector) {
    if (degree > 0) return degree;
    if (vector) return 2;
    return l == 0 ? 3 : 1;
}

// Rust's {:.2e} (multigrid.rs:730-733): 2 decimals, the exponent without sign padding or zeros.
inline std::string rust_e2(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.2e", v);
    std::string t(buf);
    const size_t e = t.find('e');
    if (e == std::string::npos) return t;
    return t.substr(0, e) + "e" + std::to_string(std::atoi(t.c_str() + e + 1));
}

}  // namespace fem
