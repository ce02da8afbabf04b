// The solver of the assembled boundary (fem_solver_new_csr): the reference's Solver::new for a
// CSR matrix of the free unknowns (crates/fem/src/solve.rs:92-183) on the device.
//
//  - Direct: the dense FP64 LDL^T of the whole matrix (dense.hip) up to MULTIFRONTAL_FROM unknowns,
//    the multifrontal LDL^T (multifrontal.hip) above, as the reference's factor is its multifrontal
//    Cholesky (cholesky.rs:44-160); fem_options.direct_solver picks one.
//  - Grid multigrid: the reference's hierarchy (csr_host.cpp: coarsen and Galerkin products, the
//    level loop of multigrid.rs:84-139), the matrices and transfers of every level in CSR form on
//    the device, a Chebyshev-Jacobi smoother, the dense LDL^T on the coarsest level.
//  - p-multigrid: the caller's prolongation to the linear unknowns and its Galerkin product
//    (pmultigrid.rs:46-60), then grid multigrid below it when the caller gives the linear
//    unknowns' grid, else the dense LDL^T there.
//  - The outer PCG is the one of the box solver (multigrid.rs:692-753; kernels.hpp), in FP64 on
//    the FP64 matrix, with the V-cycle in FP32 or FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "common.hpp"
#include "csr_kernels.hpp"
#include "csr_solver.hpp"
#include "dense.hpp"
#include "multifrontal.hpp"

namespace fem {

namespace {

constexpr uint64_t COARSEST = 3000;       // multigrid.rs:50

int threads_for_setup() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

// A matrix on the device in sliced ELL form (csr_kernels.hpp): slices of SELL_C rows, each
// padded to its longest row and stored column-major; base[s] is slice s's first entry.
template <class V>
struct DevCsr {
    uint64_t rows = 0, cols = 0, nnz = 0, padded = 0;
    bool narrow = false;  // 16-bit column offsets from each slice's smallest column
    DArray<uint32_t> base, col, cfirst;
    DArray<uint16_t> col16;
    DArray<V> val;

    void upload(uint64_t r, uint64_t c, const std::vector<uint64_t>& p, const std::vector<uint32_t>& ix,
                const std::vector<double>& v, hipStream_t s) {
        rows = r;
        cols = c;
        nnz = p[r];
        const uint64_t slices = (r + SELL_C - 1) / SELL_C;
        std::vector<uint32_t> b(slices + 1, 0), lo(slices, 0);
        uint64_t total = 0;
        narrow = true;
        for (uint64_t sl = 0; sl < slices; ++sl) {
            uint64_t w = 0;
            uint32_t cmin = UINT32_MAX, cmax = 0;
            for (uint64_t q = sl * SELL_C; q < std::min<uint64_t>(r, (sl + 1) * SELL_C); ++q) {
                w = std::max(w, p[q + 1] - p[q]);
                for (uint64_t k = p[q]; k < p[q + 1]; ++k) {
                    cmin = std::min(cmin, ix[k]);
                    cmax = std::max(cmax, ix[k]);
                }
            }
            lo[sl] = cmin == UINT32_MAX ? 0u : cmin;  // a slice without entries: column 0
            if (cmin != UINT32_MAX && cmax - cmin > 0xFFFFu) narrow = false;
            total += w * SELL_C;
            if (total > 0xFFFFFFFFull) throw Error("a matrix of more than 2^32 - 1 entries is beyond the device path");
            b[sl + 1] = (uint32_t)total;
        }
        padded = total;
        // padding: value 0 at the slice's smallest column
        std::vector<uint32_t> cc(std::max<uint64_t>(total, 1), 0);
        std::vector<V> vv(std::max<uint64_t>(total, 1), V(0));
        for (uint64_t sl = 0; sl < slices; ++sl) {
            const uint64_t w = (b[sl + 1] - b[sl]) / SELL_C;
            for (int lane = 0; lane < SELL_C; ++lane) {
                const uint64_t q = sl * SELL_C + lane;
                for (uint64_t jj = 0; jj < w; ++jj) {
                    const uint64_t at = b[sl] + jj * SELL_C + lane;
                    if (q < r && p[q] + jj < p[q + 1]) {
                        cc[at] = ix[p[q] + jj];
                        vv[at] = (V)v[p[q] + jj];
                    } else {
                        cc[at] = lo[sl];
                    }
                }
            }
        }
        base.resize(slices + 1);
        base.upload(b.data(), slices + 1, s);
        if (narrow) {
            std::vector<uint16_t> c16(cc.size());
            for (uint64_t sl = 0; sl < slices; ++sl)
                for (uint64_t at = b[sl]; at < b[sl + 1]; ++at) c16[at] = (uint16_t)(cc[at] - lo[sl]);
            col16.resize(c16.size());
            col16.upload(c16.data(), c16.size(), s);
            cfirst.resize(slices);
            cfirst.upload(lo.data(), slices, s);
        } else {
            col.resize(cc.size());
            col.upload(cc.data(), cc.size(), s);
        }
        val.resize(vv.size());
        val.upload(vv.data(), vv.size(), s);
        FEM_HIP(hipStreamSynchronize(s));  // the host staging goes out of scope
    }

    // f(SellMat<V, uint16_t>) or f(SellMat<V, uint32_t>), as the columns are stored.
    template <class F>
    void visit(F&& f) const {
        if (narrow)
            f(SellMat<V, uint16_t>{base.data(), cfirst.data(), col16.data(), val.data()});
        else
            f(SellMat<V, uint32_t>{base.data(), nullptr, col.data(), val.data()});
    }
};

template <class T>
struct CLevel {
    uint64_t n = 0;
    DevCsr<T> a;     // this level's matrix
    DevCsr<T> p, r;  // from the next coarser level: P (n rows) and R = P^T (coarse rows)
    DArray<T> invd, b, xa, xb, d, res;
    int cur = 0;
    double lmax = 0.0;
    std::vector<T> c1, c2;
    T* x() { return cur ? xb.data() : xa.data(); }
    T* x2() { return cur ? xa.data() : xb.data(); }
};

template <class T>
struct CHierarchy {
    std::vector<CLevel<T>> lv;
    DenseCoarse coarse;

    void cycle(size_t l, hipStream_t s, bool first_done) {
        CLevel<T>& L = lv[l];
        L.cur = 0;
        if (l + 1 == lv.size()) {
            // first_done on the dense level: the restriction wrote its dense right side
            if (first_done) coarse.solve_rhs<T>(L.x(), s);
            else coarse.solve<T>(L.b.data(), L.x(), s);
            return;
        }
        const uint64_t n = L.n;
        if (!first_done) launch(k_cheb_first<T, T>, n, s, n, (const T*)L.b.data(), (const T*)L.invd.data(), L.x(),
                                L.d.data(), L.c2[0]);
        const int deg = (int)L.c1.size();  // the level's smoothing steps
        for (int st = 1; st < deg; ++st) {
            step(L, L.c1[st], L.c2[st], s);
            L.cur ^= 1;
        }
        L.a.visit([&](auto m) {
            launch_blocks((k_csr_residual<T, T, typename decltype(m)::col_t>), nblocks(n), s, n, m, (const T*)L.x(),
                          (const T*)L.b.data(), L.res.data());
        });
        CLevel<T>& C = lv[l + 1];
        const bool smoothed = l + 2 < lv.size();
        // into the dense level: straight to its dense right side (its unknowns are in dense order)
        T* bc = smoothed ? C.b.data() : coarse.rhs<T>();
        L.r.visit([&](auto m) {
            launch_blocks((k_csr_restrict<T, T, typename decltype(m)::col_t, T>), nblocks(C.n), s, C.n, m,
                          (const T*)L.res.data(), bc, smoothed ? (const T*)C.invd.data() : (const T*)nullptr,
                          smoothed ? C.c2[0] : T(0), C.xa.data(), C.d.data());
        });
        cycle(l + 1, s, true);
        L.p.visit([&](auto m) {
            launch_blocks((k_csr_prolong_add<T, T, typename decltype(m)::col_t>), nblocks(n), s, n, m, (const T*)C.x(),
                          L.x());
        });
        for (int st = 0; st < deg; ++st) {
            step(L, st == 0 ? T(0) : L.c1[st], L.c2[st], s);
            L.cur ^= 1;
        }
    }

    void step(CLevel<T>& L, T c1, T c2, hipStream_t s) {
        L.a.visit([&](auto m) {
            launch_blocks((k_csr_cheb<T, T, typename decltype(m)::col_t, T>), nblocks(L.n), s, L.n, m, (const T*)L.x(),
                          L.x2(), L.d.data(), (const T*)L.b.data(), (const T*)L.invd.data(), c1, c2);
        });
    }
};

}  // namespace

class CsrSolverImpl {
   public:
    CsrSolverImpl(csr::Csr a, CsrUnknowns u, const Options& o) : opt_(o) {
        const auto t0 = std::chrono::steady_clock::now();
        csr::validate(a);
        const uint64_t n = a.n;
        FEM_HIP(hipSetDevice(o.device));
        hipDeviceProp_t prop;
        FEM_HIP(hipGetDeviceProperties(&prop, o.device));
        device_name_ = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ")";
        FEM_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        FEM_HIP(hipHostMalloc((void**)&pinned_, 16 * sizeof(double), hipHostMallocDefault));
        FEM_HIP(hipEventCreate(&ev0_));
        FEM_HIP(hipEventCreate(&ev1_));
        // solve.rs:98-183: the direct solver when forced, or as the crossover chooses
        if (u.cells && u.kind != FEM_UNKNOWNS_GRID)
            throw Error("cells give the coarse matrices of grid unknowns (a displacement on the box grid)");
        if (u.kind == FEM_UNKNOWNS_POINTS) {
            if (o.method == FEM_METHOD_MULTIGRID)
                throw Error("multigrid needs the box grid: a mesh of tetrahedra has the direct solver");
            direct_ = true;
        } else if (u.kind == FEM_UNKNOWNS_GRID || u.kind == FEM_UNKNOWNS_QUADRATIC) {
            if (o.method == FEM_METHOD_AUTO) {
                if (u.kind == FEM_UNKNOWNS_GRID) {
                    csr::validate(u.grid, n);
                    u.xyz = grid_xyz(u.grid);
                }
                direct_ = auto_picks_direct(a, u.xyz, o, counts_, analysed_);
            } else {
                direct_ = o.method == FEM_METHOD_DIRECT;
            }
        } else {
            throw Error("unknown kind of unknowns");
        }
        n_ = n;
        a64_.upload(n, n, a.ptr, a.idx, a.val, s_);
        for (DArray<double>* v : {&x_, &b_, &r_, &p_, &q_, &z_}) v->resize(n);
        parts_ = (n + BLOCK - 1) / BLOCK + 1;
        partA_.resize(parts_);
        partB_.resize(parts_);
        scal_.resize(1);
        red_.resize(4);
        if (direct_ && o.direct_solver != FEM_DIRECT_DENSE &&
            (o.direct_solver == FEM_DIRECT_MULTIFRONTAL || n > MULTIFRONTAL_FROM)) {
            // the multifrontal factor orders by nested dissection, which needs the unknowns' points
            if (u.kind == FEM_UNKNOWNS_GRID && u.xyz.empty()) {
                csr::validate(u.grid, n);
                u.xyz = grid_xyz(u.grid);
            }
            if (u.xyz.size() == n) {
                size_t free_b = 0, total_b = 0;
                FEM_HIP(hipMemGetInfo(&free_b, &total_b));
                mf_ = std::make_unique<Multifrontal>(a, u.xyz, s_, (uint64_t)(0.8 * (double)free_b));
            } else if (o.direct_solver == FEM_DIRECT_MULTIFRONTAL) {
                throw Error("the multifrontal factor orders by nested dissection: it needs the unknowns' coordinates");
            }
        }
        if (direct_ && !mf_) {
            if (n > DENSE_DIRECT_LIMIT) {
                std::string why;
                if (analysed_)
                    why = " (the crossover chose it: the factor has " + std::to_string(counts_.entries) + " entries and " +
                          rust_e(counts_.flops) + " flops; ask for multigrid)";
                throw Error("the direct solve holds at most " + std::to_string(DENSE_DIRECT_LIMIT) +
                            " unknowns; this matrix has " + std::to_string(n) + why);
            }
            factor_dense(a, direct_coarse_, false);
        } else {
            build_hierarchy(std::move(a), std::move(u));
        }
        FEM_HIP(hipStreamSynchronize(s_));
        setup_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    ~CsrSolverImpl() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        if (pinned_) (void)hipHostFree(pinned_);
        if (ev0_) (void)hipEventDestroy(ev0_);
        if (ev1_) (void)hipEventDestroy(ev1_);
        if (s_) (void)hipStreamDestroy(s_);
    }

    int method() const { return direct_ ? FEM_METHOD_DIRECT : FEM_METHOD_MULTIGRID; }
    uint64_t n() const { return n_; }
    bool crossover_counts(uint64_t* entries, double* flops) const {
        if (!analysed_) return false;
        if (entries) *entries = counts_.entries;
        if (flops) *flops = counts_.flops;
        return true;
    }
    Options& options() { return opt_; }
    const char* device_name() const { return device_name_.c_str(); }
    const csr::GridHierarchy& host_levels() const { return host_; }

    void level_info(int l, uint64_t* unknowns, double* lmax) const {
        if (direct_) {
            if (l != 0) throw Error("no such level");
            *unknowns = n_;
            *lmax = 0.0;
            return;
        }
        const size_t count = h32_ ? h32_->lv.size() : h64_->lv.size();
        if (l < 0 || (size_t)l >= count) throw Error("no such level");
        *unknowns = h32_ ? h32_->lv[l].n : h64_->lv[l].n;
        *lmax = h32_ ? h32_->lv[l].lmax : h64_->lv[l].lmax;
    }

    void apply_host(const double* x, double* y) {
        x_.upload(x, n_, s_);
        apply64(x_.data(), q_.data());
        q_.download(y, n_, s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    void precondition_host(const double* r, double* z) {
        if (direct_) throw Error("the direct solver has no preconditioner");
        r_.upload(r, n_, s_);
        precondition();
        z_.download(z, n_, s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    void solve_host(const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* st) {
        DArray<double> db(n_), dx(n_);
        db.upload(b, n_, s_);
        dx.upload(x, n_, s_);
        std::string error;
        try {
            solve_device(db.data(), dx.data(), cancelled, ctx, st);
        } catch (const Error& e) {
            error = e.what();
        }
        dx.download(x, n_, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        if (!error.empty()) throw Error(error);
    }

    void solve_device(const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx, fem_stats* st) {
        fem_stats out{};
        out.method = method();
        out.precision = opt_.precision;
        out.setup_seconds = setup_seconds_;
        FEM_HIP(hipMemcpyAsync(b_.data(), d_b, n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        FEM_HIP(hipMemcpyAsync(x_.data(), d_x, n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        FEM_HIP(hipEventRecord(ev0_, s_));
        std::string error;
        try {
            if (direct_) {
                // solve.rs:207-213: x = A^-1 b, the true relative residual
                if (mf_) mf_->solve(b_.data(), x_.data(), s_);
                else direct_coarse_.solve<double>(b_.data(), x_.data(), s_);
                out.iterations = 0;
                out.levels = 1;
                out.coarse_unknowns = n_;
                out.residual = out.true_residual = true_residual();
            } else {
                pcg(cancelled, ctx, out);
            }
        } catch (const Error& e) {
            error = e.what();
        }
        FEM_HIP(hipEventRecord(ev1_, s_));
        FEM_HIP(hipMemcpyAsync(d_x, x_.data(), n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        float ms = 0.0f;
        FEM_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        out.seconds = ms * 1e-3;
        if (st) *st = out;
        if (!error.empty()) throw Error(error);
    }

   private:
    // ---------------------------------------------------------------- setup
    // The dense matrix of a CSR on the device (rows scattered by a kernel, so no n^2 host array),
    // then its factor and inverse.
    void factor_dense(const csr::Csr& a, DenseCoarse& dc, bool fp32) {
        const uint64_t n = a.n;
        DArray<uint64_t> dptr(n + 1);
        DArray<uint32_t> didx(std::max<uint64_t>(a.ptr[n], 1));
        DArray<double> dval(std::max<uint64_t>(a.ptr[n], 1));
        dptr.upload(a.ptr.data(), n + 1, s_);
        if (a.ptr[n]) {
            didx.upload(a.idx.data(), a.ptr[n], s_);
            dval.upload(a.val.data(), a.ptr[n], s_);
        }
        DArray<double> A(n * n);
        launch(k_fill<double>, n * n, s_, n * n, A.data(), 0.0);
        launch(k_csr_scatter_dense, n, s_, n, (const uint64_t*)dptr.data(), (const uint32_t*)didx.data(),
               (const double*)dval.data(), A.data());
        FEM_HIP(hipStreamSynchronize(s_));
        dc.factor(A.data(), (int)n, fp32, s_);
        std::vector<int64_t> iota(n);
        std::iota(iota.begin(), iota.end(), (int64_t)0);
        dc.map().resize(n);
        dc.map().upload(iota.data(), n, s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    void build_hierarchy(csr::Csr a, CsrUnknowns u) {
        const int threads = threads_for_setup();
        const uint64_t cmax = opt_.coarse_max;
        if (u.kind == FEM_UNKNOWNS_GRID) {
            csr::validate(u.grid, a.n);
            if (u.cells) csr::validate(u.cell_data, u.grid);
            host_ = csr::grid_hierarchy(a, u.grid, COARSEST, opt_.levels, threads, u.cells ? &u.cell_data : nullptr);
        } else {
            // pmultigrid.rs:46-60: the linear level is P^T A P; the diagonal must be above 0
            csr::validate(u.p);
            if (u.p.rows != a.n) throw Error("the prolongation needs a row for each unknown");
            for (double d : csr::diagonal(a))
                if (!(d > 0.0)) throw Error("the matrix has a diagonal entry that is not above 0");
            csr::Csr ac = csr::galerkin(u.p, a, threads);
            host_.a.push_back(std::move(a));
            host_.level.emplace_back();
            if (u.coarse_grid) {
                csr::validate(u.coarse, ac.n);
                const int below = opt_.levels > 0 ? std::max(1, opt_.levels - 1) : 0;
                csr::GridHierarchy g = csr::grid_hierarchy(ac, u.coarse, COARSEST, below, threads);
                host_.p.push_back(std::move(u.p));
                for (size_t l = 0; l < g.a.size(); ++l) {
                    host_.a.push_back(std::move(g.a[l]));
                    host_.level.push_back(std::move(g.level[l]));
                    if (l < g.p.size()) host_.p.push_back(std::move(g.p[l]));
                }
            } else {
                host_.p.push_back(std::move(u.p));
                host_.a.push_back(std::move(ac));
                host_.level.emplace_back();
            }
        }
        if (host_.a.size() < 2) throw Error("the matrix is too small for multigrid: use the direct solver");
        // A quadratic caller with no grid below its linear level (the reference's tet10 caller,
        // elastic_tet.rs:294-313, passes coordinates only) has that level factored whole, as the
        // reference does (pmultigrid.rs:46-60): the dense LDL^T up to DENSE_DIRECT_LIMIT unknowns,
        // whatever coarse_max says. Below a grid hierarchy the coarsest level keeps coarse_max.
        const bool whole_linear = u.kind == FEM_UNKNOWNS_QUADRATIC && !u.coarse_grid;
        const uint64_t limit = whole_linear ? std::max<uint64_t>(cmax, DENSE_DIRECT_LIMIT) : cmax;
        if (host_.a.back().n > limit)
            throw Error("the coarsest level has " + std::to_string(host_.a.back().n) +
                        " unknowns, more than the dense limit " + std::to_string(limit));
        for (const csr::Csr& m : host_.a)
            for (double d : csr::diagonal(m))
                if (!(d > 0.0)) throw Error("the matrix has a diagonal entry that is not above 0");
        if (opt_.precision == FEM_PREC_FP64) {
            h64_ = std::make_unique<CHierarchy<double>>();
            make_levels(*h64_);
        } else {
            h32_ = std::make_unique<CHierarchy<float>>();
            make_levels(*h32_);
        }
    }

    template <class T>
    void make_levels(CHierarchy<T>& H) {
        const size_t count = host_.a.size();
        H.lv.resize(count);
        for (size_t l = 0; l < count; ++l) {
            CLevel<T>& L = H.lv[l];
            const csr::Csr& m = host_.a[l];
            L.n = m.n;
            for (DArray<T>* v : {&L.b, &L.xa, &L.xb, &L.d, &L.res, &L.invd}) v->resize(L.n);
            launch(k_fill<T>, L.n, s_, L.n, L.xa.data(), T(0));
            launch(k_fill<T>, L.n, s_, L.n, L.xb.data(), T(0));
            launch(k_fill<T>, L.n, s_, L.n, L.d.data(), T(0));
            if (l + 1 == count) break;  // the dense level keeps only its vectors
            L.a.upload(m.n, m.n, m.ptr, m.idx, m.val, s_);
            const csr::Rect& P = host_.p[l];
            L.p.upload(P.rows, P.cols, P.ptr, P.idx, P.val, s_);
            const csr::Rect R = csr::transpose(P);
            L.r.upload(R.rows, R.cols, R.ptr, R.idx, R.val, s_);
            const std::vector<double> dg = csr::diagonal(m);
            std::vector<T> inv(L.n);
            for (uint64_t i = 0; i < L.n; ++i) inv[i] = (T)(1.0 / dg[i]);
            L.invd.upload(inv.data(), L.n, s_);
            FEM_HIP(hipStreamSynchronize(s_));
            L.lmax = estimate_lmax(L);
            // automatic: 2 on every level. A finest step moves the whole matrix (6-8 B per FP32
            // entry), so the matrix-free path's 3-then-1 choice costs more than the iteration it
            // saves (heat 128^3 with 32-bit columns: 168 against 197 MDOF/s).
            chebyshev(L, opt_.degree > 0 ? opt_.degree : 2);
        }
        factor_dense(host_.a.back(), H.coarse, sizeof(T) == 4);
    }

    template <class T>
    double estimate_lmax(CLevel<T>& L) {
        const uint64_t n = L.n;
        T* v = L.xa.data();
        T* w = L.xb.data();
        launch(k_random<T>, n, s_, n, v, (const T*)L.invd.data());
        const unsigned pb = std::min<unsigned>((unsigned)((n + BLOCK - 1) / BLOCK), 1024);
        for (int it = 0; it < opt_.power_steps; ++it) {
            launch_blocks(k_dot<T>, pb, s_, n, (const T*)v, (const T*)v, partA_.data());
            launch_reduce(k_sum, s_, (const double*)partA_.data(), (int)pb, red_.data());
            launch(k_normalise<T>, n, s_, n, v, (const double*)red_.data());
            L.a.visit([&](auto m) {
                launch_blocks((k_csr_apply<T, T, typename decltype(m)::col_t>), nblocks(n), s_, n, m, (const T*)v, w);
            });
            launch(k_scale<T>, n, s_, n, w, (const T*)L.invd.data());
            if (it + 1 == opt_.power_steps) {
                launch_blocks(k_dot<T>, pb, s_, n, (const T*)w, (const T*)w, partA_.data());
                launch_reduce(k_sum, s_, (const double*)partA_.data(), (int)pb, red_.data() + 1);
            }
            std::swap(v, w);
        }
        double nw = 0.0;
        if (opt_.power_steps > 0) {
            FEM_HIP(hipMemcpyAsync(&nw, red_.data() + 1, sizeof(double), hipMemcpyDeviceToHost, s_));
            FEM_HIP(hipStreamSynchronize(s_));
        }
        launch(k_fill<T>, n, s_, n, L.xa.data(), T(0));
        launch(k_fill<T>, n, s_, n, L.xb.data(), T(0));
        return std::sqrt(nw);
    }

    // The Chebyshev recurrence on [cheb_lower, cheb_upper] lmax (as the box solver's).
    template <class T>
    void chebyshev(CLevel<T>& L, int degree) {
        const double hi = opt_.cheb_upper * L.lmax, lo = opt_.cheb_lower * L.lmax;
        const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
        L.c1.assign(degree, T(0));
        L.c2.assign(degree, T(0));
        double rho = 1.0 / sigma;
        L.c2[0] = T(1.0 / theta);
        for (int s = 1; s < degree; ++s) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            L.c1[s] = T(rn * rho);
            L.c2[s] = T(2.0 * rn / delta);
            rho = rn;
        }
    }

    // ---------------------------------------------------------------- the operator and PCG
    void apply64(const double* x, double* y) {
        a64_.visit([&](auto m) {
            launch_blocks((k_csr_apply<double, double, typename decltype(m)::col_t>), nblocks(n_), s_, n_, m, x, y);
        });
    }

    // |b - A x| / |b| in FP64.
    double true_residual() {
        const unsigned nb = start_residual();
        launch_reduce(k_sum, s_, (const double*)partA_.data(), (int)nb, red_.data());
        launch_reduce(k_sum, s_, (const double*)partB_.data(), (int)nb, red_.data() + 1);
        double v[2];
        FEM_HIP(hipMemcpyAsync(v, red_.data(), 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        const double bn = std::sqrt(v[0]), rn = std::sqrt(v[1]);
        return bn > 0.0 ? rn / bn : rn;
    }

    // r = b - A x with the partials of b.b (partA_) and r.r (partB_); returns their count.
    unsigned start_residual() {
        const unsigned nb = nblocks(n_);
        a64_.visit([&](auto m) {
            launch_blocks((k_csr_pcg_start<typename decltype(m)::col_t>), nb, s_, n_, m, (const double*)x_.data(),
                          (const double*)b_.data(), r_.data(), partA_.data(), partB_.data());
        });
        return nb;
    }

    // z = M^-1 r (the V-cycle from its own first step).
    void precondition() {
        if (h32_) {
            CLevel<float>& L = h32_->lv[0];
            launch((k_convert<double, float>), n_, s_, n_, (const double*)r_.data(), L.b.data());
            h32_->cycle(0, s_, false);
            launch((k_convert<float, double>), n_, s_, n_, (const float*)L.x(), z_.data());
        } else {
            CLevel<double>& L = h64_->lv[0];
            FEM_HIP(hipMemcpyAsync(L.b.data(), r_.data(), n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
            h64_->cycle(0, s_, false);
            FEM_HIP(hipMemcpyAsync(z_.data(), L.x(), n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        }
    }

    template <class T>
    void update_and_vcycle(CHierarchy<T>& H) {
        CLevel<T>& L = H.lv[0];
        const bool smooths = H.lv.size() > 1;
        const unsigned nbv = rblocks(n_);
        launch_blocks((k_pcg_update<T, T>), nbv, s_, n_, (const PcgScalars*)scal_.data(), r_.data(),
                      (const double*)q_.data(), L.b.data(), partB_.data(),
                      smooths ? (const T*)L.invd.data() : (const T*)nullptr, smooths ? L.c2[0] : T(0), L.xa.data(),
                      L.d.data());
        H.cycle(0, s_, smooths);  // r.r (partB_) is summed with r.z after it
    }

    // One PCG iteration after the head (multigrid.rs:738-750), gated on the device (PcgScalars).
    void iteration_body() {
        const unsigned nbv = rblocks(n_);
        PcgScalars* sc = scal_.data();
        const unsigned nq = nblocks(n_);
        a64_.visit([&](auto m) {
            launch_blocks((k_csr_pcg_q<typename decltype(m)::col_t>), nq, s_, n_, m, (const double*)p_.data(), q_.data(),
                          partA_.data());
        });
        launch_reduce(k_pcg_scalars, s_, sc, (const double*)partA_.data(), (const double*)nullptr, (int)nq, (int)nq, 1);
        if (h32_) {
            update_and_vcycle(*h32_);
            launch_blocks((k_pcg_rz<float>), nbv, s_, n_, (const float*)h32_->lv[0].x(), (double*)nullptr,
                          (const double*)r_.data(), partA_.data());
        } else {
            update_and_vcycle(*h64_);
            launch_blocks((k_pcg_rz<double>), nbv, s_, n_, (const double*)h64_->lv[0].x(), (double*)nullptr,
                          (const double*)r_.data(), partA_.data());
        }
        launch_reduce(k_pcg_scalars, s_, sc, (const double*)partA_.data(), (const double*)partB_.data(), (int)nbv, (int)nbv, 5);
        if (h32_)
            launch((k_pcg_dir<float>), n_, s_, n_, (const PcgScalars*)sc, (const float*)h32_->lv[0].x(), p_.data(),
                   x_.data(), 0);
        else
            launch((k_pcg_dir<double>), n_, s_, n_, (const PcgScalars*)sc, (const double*)h64_->lv[0].x(), p_.data(),
                   x_.data(), 0);
    }

    void iterate() {
        if (!graph_) {
            hipGraph_t g = nullptr;
            FEM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
            iteration_body();
            FEM_HIP(hipStreamEndCapture(s_, &g));
            FEM_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            FEM_HIP(hipGraphDestroy(g));
        }
        FEM_HIP(hipGraphLaunch(graph_, s_));
    }

    const PcgScalars& read_scalars() {
        FEM_HIP(hipMemcpyAsync(pinned_, scal_.data(), sizeof(PcgScalars), hipMemcpyDeviceToHost, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        return *reinterpret_cast<const PcgScalars*>(pinned_);
    }

    // multigrid.rs:692-753 with the refinement restarts of the box solver: passes of PCG from the
    // current x, each ended by the device's stopping test, queued in batches between readbacks.
    void pcg(int (*cancelled)(void*), void* ctx, fem_stats& out) {
        const unsigned nbv = rblocks(n_);
        uint64_t total = 0;
        int restarts = 0;
        for (;;) {
            const double limit = (double)(opt_.max_iter - std::min<uint64_t>(opt_.max_iter, total));
            hipLaunchKernelGGL(k_pcg_config, dim3(1), dim3(64), 0, s_, scal_.data(), opt_.tol, limit);
            const unsigned nb = start_residual();
            launch_reduce(k_pcg_scalars, s_, scal_.data(), (const double*)partA_.data(), (const double*)partB_.data(),
                          (int)nb, (int)nb, 0);
            PcgScalars ps = read_scalars();
            const double norm_b = std::sqrt(ps.bb);
            if (norm_b == 0.0) {  // multigrid.rs:706-710
                launch(k_fill<double>, n_, s_, n_, x_.data(), 0.0);
                out.iterations = total;
                out.residual = out.true_residual = 0.0;
                out.levels = h32_ ? h32_->lv.size() : h64_->lv.size();
                out.coarse_unknowns = host_.a.back().n;
                out.restarts = restarts;
                return;
            }
            precondition();
            launch_blocks((k_pcg_rz<float>), nbv, s_, n_, (const float*)nullptr, z_.data(), (const double*)r_.data(),
                          partA_.data());
            launch_reduce(k_pcg_scalars, s_, scal_.data(), (const double*)partA_.data(), (const double*)nullptr,
                          (int)nbv, (int)nbv, 3);
            launch((k_pcg_dir<double>), n_, s_, n_, (const PcgScalars*)scal_.data(), (const double*)z_.data(), p_.data(),
                   x_.data(), 1);
            uint64_t it = 0;
            double residual = std::sqrt(ps.rr) / norm_b, rate = 0.0;
            bool stopped = ps.active == 0.0;
            for (;;) {
                const bool at_limit = total + it >= opt_.max_iter;
                if (residual <= opt_.tol || (stopped && !at_limit)) break;
                if (at_limit) {
                    out.iterations = total + it;
                    out.residual = residual;
                    out.restarts = restarts;
                    throw Error("conjugate gradients did not converge: the residual is " + rust_e2(residual) + " after " +
                                std::to_string(opt_.max_iter) + " iterations");
                }
                if (cancelled && cancelled(ctx)) throw Error("cancelled");
                // half of the iterations the last batch's rate predicts, at most 8
                int k = 1;
                if (rate > 0.0 && rate < 1.0) {
                    const double need = std::ceil(std::log(opt_.tol / residual) / std::log(rate));
                    const double left = (double)(opt_.max_iter - (total + it));
                    k = (int)std::max(1.0, std::min({std::floor(0.5 * need), 8.0, left}));
                }
                for (int q = 0; q < k; ++q) iterate();
                ps = read_scalars();
                const uint64_t it2 = (uint64_t)ps.it;
                const double res2 = std::sqrt(ps.rr) / norm_b;
                if (it2 > it && residual > 0.0 && res2 > 0.0) rate = std::pow(res2 / residual, 1.0 / (double)(it2 - it));
                it = it2;
                residual = res2;
                stopped = ps.active == 0.0;
            }
            total += it;
            const double tr = true_residual();
            out.iterations = total;
            out.residual = residual;
            out.true_residual = tr;
            out.restarts = restarts;
            out.levels = h32_ ? h32_->lv.size() : h64_->lv.size();
            out.coarse_unknowns = host_.a.back().n;
            if (!opt_.refine || tr <= opt_.tol || restarts >= 5) return;
            ++restarts;
        }
    }

    Options opt_;
    std::string device_name_;
    hipStream_t s_ = nullptr;
    double* pinned_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    bool direct_ = false;
    bool analysed_ = false;  // a FLOPS crossover analysed the matrix (counts_)
    sym::Counts counts_;
    uint64_t n_ = 0, parts_ = 0;
    double setup_seconds_ = 0.0;
    DevCsr<double> a64_;
    DArray<double> x_, b_, r_, p_, q_, z_, partA_, partB_, red_;
    DArray<PcgScalars> scal_;
    DenseCoarse direct_coarse_;
    std::unique_ptr<Multifrontal> mf_;
    csr::GridHierarchy host_;
    std::unique_ptr<CHierarchy<float>> h32_;
    std::unique_ptr<CHierarchy<double>> h64_;
};

CsrSolver::CsrSolver(csr::Csr a, CsrUnknowns u, const Options& o)
    : impl_(std::make_unique<CsrSolverImpl>(std::move(a), std::move(u), o)) {}
CsrSolver::~CsrSolver() = default;
int CsrSolver::method() const { return impl_->method(); }
uint64_t CsrSolver::n() const { return impl_->n(); }
Options& CsrSolver::options() { return impl_->options(); }
void CsrSolver::solve_device(const double* d_b, double* d_x, int (*c)(void*), void* ctx, fem_stats* st) {
    impl_->solve_device(d_b, d_x, c, ctx, st);
}
void CsrSolver::solve_host(const double* b, double* x, int (*c)(void*), void* ctx, fem_stats* st) {
    impl_->solve_host(b, x, c, ctx, st);
}
void CsrSolver::apply_host(const double* x, double* y) { impl_->apply_host(x, y); }
void CsrSolver::precondition_host(const double* r, double* z) { impl_->precondition_host(r, z); }
void CsrSolver::level_info(int l, uint64_t* n, double* lmax) const { impl_->level_info(l, n, lmax); }
const char* CsrSolver::device_name() const { return impl_->device_name(); }
bool CsrSolver::crossover_counts(uint64_t* e, double* f) const { return impl_->crossover_counts(e, f); }

std::vector<std::array<double, 3>> grid_xyz(const csr::GridLevel& g) {
    std::vector<std::array<double, 3>> xyz(g.points.size());
    for (size_t u = 0; u < g.points.size(); ++u)
        for (int d = 0; d < 3; ++d) xyz[u][d] = g.planes[d][g.points[u][d]];
    return xyz;
}

bool auto_picks_direct(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, const Options& o,
                       sym::Counts& counts, bool& analysed) {
    analysed = false;
    if (o.crossover_kind == FEM_CROSSOVER_UNKNOWNS) return a.n <= o.crossover;
    if (o.crossover_kind != FEM_CROSSOVER_FLOPS) throw Error("unknown crossover");
    if (xyz.size() != a.n) throw Error("the FLOPS crossover needs the coordinates of each unknown");
    counts = sym::analyse(a, sym::nested_dissection(a, xyz));
    analysed = true;
    return sym::flops_picks_direct(counts, a.n, o.crossover_flops, o.crossover_entries, o.crossover_flat);
}
const csr::GridHierarchy& CsrSolver::host_levels() const { return impl_->host_levels(); }

}  // namespace fem

// ---------------------------------------------------------------------------------------------
// C ABI pieces of the assembled boundary (the solver handle itself lives in solver.hip).
namespace fem {

csr::Csr csr_from(const fem_csr* a) {
    if (!a || !a->ptr || (a->rows && (!a->idx || !a->val) && a->ptr[a->rows])) throw Error("no matrix");
    if (a->cols != a->rows) throw Error("the matrix must be square");
    csr::Csr m;
    m.n = a->rows;
    m.ptr.assign(a->ptr, a->ptr + a->rows + 1);
    const uint64_t nnz = m.ptr[m.n];
    m.idx.assign(a->idx, a->idx + nnz);
    m.val.assign(a->val, a->val + nnz);
    return m;
}

csr::Rect rect_from(const fem_csr* p) {
    if (!p || !p->ptr) throw Error("no prolongation");
    csr::Rect r;
    r.rows = p->rows;
    r.cols = p->cols;
    r.ptr.assign(p->ptr, p->ptr + p->rows + 1);
    const uint64_t nnz = r.ptr[r.rows];
    if (nnz && (!p->idx || !p->val)) throw Error("no prolongation entries");
    r.idx.assign(p->idx, p->idx + nnz);
    r.val.assign(p->val, p->val + nnz);
    return r;
}

csr::GridLevel grid_from(const fem_grid_level* g, uint64_t n) {
    if (!g || !g->points || !g->comp) throw Error("no grid level");
    csr::GridLevel l;
    for (int d = 0; d < 3; ++d) {
        if (!g->planes[d] || g->nplanes[d] == 0) throw Error("each axis of the grid needs a plane");
        l.planes[d].assign(g->planes[d], g->planes[d] + g->nplanes[d]);
    }
    l.points.resize(n);
    for (uint64_t u = 0; u < n; ++u)
        for (int d = 0; d < 3; ++d) l.points[u][d] = g->points[3 * u + d];
    l.comp.assign(g->comp, g->comp + n);
    return l;
}

csr::Cells cells_from(const fem_cells* c) {
    if (!c || (c->count && (!c->origin || !c->element))) throw Error("the cells need their origins and matrices");
    csr::Cells out;
    out.origin.resize(c->count);
    out.k.resize(c->count);
    for (uint64_t e = 0; e < c->count; ++e) {
        for (int d = 0; d < 3; ++d) out.origin[e][d] = c->origin[3 * e + d];
        std::copy(c->element + 576 * e, c->element + 576 * (e + 1), out.k[e].begin());
    }
    return out;
}

CsrUnknowns unknowns_from(const fem_unknowns* u, uint64_t n) {
    if (!u) throw Error("no unknowns");
    CsrUnknowns out;
    out.kind = u->kind;
    if (u->kind == FEM_UNKNOWNS_GRID) {
        out.grid = grid_from(&u->grid, n);
    } else if (u->kind == FEM_UNKNOWNS_QUADRATIC) {
        out.p = rect_from(&u->p);
        if (u->coarse_has_grid) {
            out.coarse_grid = true;
            out.coarse = grid_from(&u->coarse_grid, out.p.cols);
        }
    } else if (u->kind != FEM_UNKNOWNS_POINTS) {
        throw Error("unknown kind of unknowns");
    }
    if (u->xyz && u->kind != FEM_UNKNOWNS_GRID) {
        out.xyz.resize(n);
        for (uint64_t k = 0; k < n; ++k)
            for (int d = 0; d < 3; ++d) out.xyz[k][d] = u->xyz[3 * k + d];
    }
    out.cells = u->cells != nullptr;
    if (u->cells) out.cell_data = cells_from(u->cells);
    return out;
}

}  // namespace fem

struct fem_csr_levels {
    fem::csr::GridHierarchy h;
};

namespace {
void put_error(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}
}  // namespace

extern "C" {

int fem_test_csr_levels_new(const fem_csr* a, const fem_grid_level* g, const fem_cells* cells, uint64_t coarsest,
                            int max_levels, fem_csr_levels** out, char* err, size_t errlen) {
    try {
        if (!out) throw fem::Error("an output handle is needed");
        *out = nullptr;
        fem::csr::Csr m = fem::csr_from(a);
        fem::csr::validate(m);
        fem::csr::GridLevel l = fem::grid_from(g, m.n);
        fem::csr::validate(l, m.n);
        fem::csr::Cells cl;
        if (cells) {
            cl = fem::cells_from(cells);
            fem::csr::validate(cl, l);
        }
        auto* h = new fem_csr_levels;
        h->h = fem::csr::grid_hierarchy(m, l, coarsest, max_levels, 8, cells ? &cl : nullptr);
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        put_error(err, errlen, e.what());
        return 1;
    }
}

int fem_test_csr_levels_count(const fem_csr_levels* h) { return h ? (int)h->h.a.size() : 0; }

int fem_test_csr_level_sizes(const fem_csr_levels* h, int l, uint64_t* n, uint64_t* a_nnz, uint64_t* p_cols,
                             uint64_t* p_nnz) {
    if (!h || l < 0 || (size_t)l >= h->h.a.size()) return 1;
    const auto& a = h->h.a[l];
    if (n) *n = a.n;
    if (a_nnz) *a_nnz = a.nnz();
    const bool has_p = (size_t)l < h->h.p.size();
    if (p_cols) *p_cols = has_p ? h->h.p[l].cols : 0;
    if (p_nnz) *p_nnz = has_p ? h->h.p[l].ptr[h->h.p[l].rows] : 0;
    return 0;
}

int fem_test_csr_level_arrays(const fem_csr_levels* h, int l, uint64_t* a_ptr, uint32_t* a_idx, double* a_val,
                              uint64_t* p_ptr, uint32_t* p_idx, double* p_val, uint32_t* points, uint8_t* comp) {
    if (!h || l < 0 || (size_t)l >= h->h.a.size()) return 1;
    const auto& a = h->h.a[l];
    if (a_ptr) std::memcpy(a_ptr, a.ptr.data(), a.ptr.size() * sizeof(uint64_t));
    if (a_idx) std::memcpy(a_idx, a.idx.data(), a.idx.size() * sizeof(uint32_t));
    if (a_val) std::memcpy(a_val, a.val.data(), a.val.size() * sizeof(double));
    if ((size_t)l < h->h.p.size()) {
        const auto& p = h->h.p[l];
        if (p_ptr) std::memcpy(p_ptr, p.ptr.data(), p.ptr.size() * sizeof(uint64_t));
        if (p_idx) std::memcpy(p_idx, p.idx.data(), p.idx.size() * sizeof(uint32_t));
        if (p_val) std::memcpy(p_val, p.val.data(), p.val.size() * sizeof(double));
    }
    const auto& g = h->h.level[l];
    for (size_t u = 0; u < g.points.size(); ++u) {
        if (points)
            for (int d = 0; d < 3; ++d) points[3 * u + d] = g.points[u][d];
        if (comp) comp[u] = g.comp[u];
    }
    return 0;
}

void fem_test_csr_levels_free(fem_csr_levels* h) { delete h; }

int fem_test_symbolic(const fem_csr* a, const double* xyz, uint64_t* order_out, uint64_t* entries, double* flops,
                      uint64_t* largest_front, uint64_t* supernodes, char* err, size_t errlen) {
    try {
        fem::csr::Csr m = fem::csr_from(a);
        fem::csr::validate(m);
        if (!xyz) throw fem::Error("the coordinates are needed");
        std::vector<std::array<double, 3>> p(m.n);
        for (uint64_t k = 0; k < m.n; ++k)
            for (int d = 0; d < 3; ++d) p[k][d] = xyz[3 * k + d];
        const std::vector<uint64_t> order = fem::sym::nested_dissection(m, p);
        const fem::sym::Counts c = fem::sym::analyse(m, order);
        if (order_out) std::copy(order.begin(), order.end(), order_out);
        if (entries) *entries = c.entries;
        if (flops) *flops = c.flops;
        if (largest_front) *largest_front = c.largest_front;
        if (supernodes) *supernodes = c.supernodes;
        return 0;
    } catch (const std::exception& e) {
        put_error(err, errlen, e.what());
        return 1;
    }
}

int fem_test_supernodal(const fem_csr* a, const double* xyz, uint64_t join, uint64_t* nsup, uint64_t* nrows,
                        uint64_t* perm, uint64_t* start, uint64_t* rptr, uint64_t* rows, int64_t* sparent, char* err,
                        size_t errlen) {
    try {
        fem::csr::Csr m = fem::csr_from(a);
        fem::csr::validate(m);
        if (!xyz) throw fem::Error("the coordinates are needed");
        std::vector<std::array<double, 3>> p(m.n);
        for (uint64_t k = 0; k < m.n; ++k)
            for (int d = 0; d < 3; ++d) p[k][d] = xyz[3 * k + d];
        const fem::sym::Supernodal S =
            fem::sym::amalgamate(fem::sym::supernodal(m, fem::sym::nested_dissection(m, p)), join);
        if (nsup) *nsup = S.supernodes();
        if (nrows) *nrows = S.rows.size();
        if (perm) std::copy(S.perm.begin(), S.perm.end(), perm);
        if (start) std::copy(S.start.begin(), S.start.end(), start);
        if (rptr) std::copy(S.rptr.begin(), S.rptr.end(), rptr);
        if (rows) std::copy(S.rows.begin(), S.rows.end(), rows);
        if (sparent) std::copy(S.sparent.begin(), S.sparent.end(), sparent);
        return 0;
    } catch (const std::exception& e) {
        put_error(err, errlen, e.what());
        return 1;
    }
}

}  // extern "C"
