// The solver: the multigrid hierarchy on the device, the V-cycle, the FP64 PCG of the reference
// (crates/fem/src/multigrid.rs:692-753) and the C ABI of include/fem_solver.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/fem_solver.h"
#include "../../include/fem_testing.h"
#include "common.hpp"
#include "dense.hpp"
#include "hex_elastic.hpp"
#include "kernels.hpp"
#include "model.hpp"
#include "multifrontal.hpp"
#include "march.hpp"
#include "tet10_fused.hpp"
#include "tet10_kuhn.hpp"
#include "setup_kernels.hpp"
#include "solver_common.hpp"
#include "csr_solver.hpp"

namespace fem {

namespace {

enum class OpKind { Hex8K, Stencil27, Tet4Heat, Tet4Elastic, Tet10 };

// One level of the hierarchy in storage type T.
template <class T>
struct Level {
    OpKind op = OpKind::Hex8K;
    int comps = 1;
    GridDims g{};
    int cells[3] = {0, 0, 0};  // cells of the level (tet10: the cells of its grid, nodes 2 n + 1)
    double h[3] = {0, 0, 0};
    bool kuhn = false;         // transfer to the next level is Kuhn P1 (else trilinear)
    DArray<T> coef;
    DArray<T> invd, b, xa, xb, d, r;
    // BF16 storage (FP32 levels only): the per-cell coefficients (not the Galerkin stencils) and D^-1
    // held as BF16 values, in 16 bits for the kernels and as their FP32 values in coef and invd
    bool bf16 = false;
    DArray<bf16_t> coef16, invd16;
    // tet10: the partial sums of the fused kernels (tet10_fused.hpp, 8 slots per unknown) and the
    // nodes their boundary launch finishes, for the tile of T
    DArray<T> P;
    DArray<int> blist;
    int nbound = 0;
    int cur = 0;
    double lmax = 0.0;
    std::vector<T> c1, c2;     // Chebyshev step coefficients
    size_t nodes() const { return g.nodes(); }
    // stored bytes of D^-1 and of the operator's coefficients as the smoothing kernels read them
    size_t invd_bytes() const { return bf16 ? invd16.bytes() : invd.bytes(); }
    size_t coef_bytes() const { return (bf16 && op != OpKind::Stencil27) ? coef16.bytes() : coef.bytes(); }
    // the hex8 operator through the column-marching kernels (31-bit byte offsets)
    bool march() const { return op == OpKind::Hex8K && nodes() * sizeof(T) < MARCH_MAX_BYTES; }
    size_t dofs() const { return g.nodes() * comps; }
    size_t ncell() const { return (size_t)cells[0] * cells[1] * cells[2]; }
    T* x() { return cur ? xb.data() : xa.data(); }
    T* x2() { return cur ? xa.data() : xb.data(); }
    // The same level for another solve of the same model: the operator and smoother data
    // (coefficients, D^-1, the boundary node list) are o's (viewed: o must outlive this level),
    // the vectors and the partial-sum workspace its own (zero).
    void shadow_of(const Level& o, hipStream_t s) {
        op = o.op;
        comps = o.comps;
        g = o.g;
        for (int d = 0; d < 3; ++d) {
            cells[d] = o.cells[d];
            h[d] = o.h[d];
        }
        kuhn = o.kuhn;
        bf16 = o.bf16;
        nbound = o.nbound;
        lmax = o.lmax;
        c1 = o.c1;
        c2 = o.c2;
        cur = 0;
        coef.view(o.coef);
        invd.view(o.invd);
        coef16.view(o.coef16);
        invd16.view(o.invd16);
        blist.view(o.blist);
        b.like(o.b, s);
        xa.like(o.xa, s);
        xb.like(o.xb, s);
        d.like(o.d, s);
        r.like(o.r, s);
        P.like(o.P, s);
    }
};

template <class T>
void hex_c(const double h[3], T c[8]) {
    for (int p = 0; p < 8; ++p) {
        double v = 0.0;
        for (int d = 0; d < 3; ++d) {
            double term = 1.0;
            for (int e = 0; e < 3; ++e) {
                const bool differ = (p >> e) & 1;
                if (e == d) term *= differ ? -1.0 / h[e] : 1.0 / h[e];
                else term *= differ ? h[e] / 6.0 : h[e] / 3.0;
            }
            v += term;
        }
        c[p] = (T)v;
    }
}

// The element type of a pointer.
template <class P>
using elem_t = std::remove_const_t<std::remove_pointer_t<P>>;

// Calls f with the level's operator, whose per-cell coefficients are read in their stored type:
// T, or BF16 on a level of an FP32 hierarchy in BF16 storage (Galerkin stencils stay in T).
template <class T, class F>
void with_op(Level<T>& L, F&& f) {
    const T ih[3] = {T(1.0 / L.h[0]), T(1.0 / L.h[1]), T(1.0 / L.h[2])};
    const T vol = T(L.h[0] * L.h[1] * L.h[2] / 6.0);
    auto build = [&](auto cp) {
        using S = elem_t<decltype(cp)>;
        switch (L.op) {
            case OpKind::Hex8K: {
                Hex8K<T, S> o{L.g, cp, {}};
                hex_c<T>(L.h, o.c);
                f(ScalarOp<Hex8K<T, S>>{o});
                break;
            }
            case OpKind::Stencil27:
                f(ScalarOp<Stencil27<T>>{Stencil27<T>{L.g, L.coef.data()}});
                break;
            case OpKind::Tet4Heat:
                f(ScalarOp<Tet4Heat<T, S>>{Tet4Heat<T, S>{L.g, cp, {ih[0], ih[1], ih[2]}, vol}});
                break;
            case OpKind::Tet4Elastic:
                f(VectorOp<Tet4Elastic<T, S>>{Tet4Elastic<T, S>{L.g, cp, {ih[0], ih[1], ih[2]}, vol}});
                break;
            case OpKind::Tet10:
                throw Error("internal: the tet10 operator runs through its fused kernels");
        }
    };
    if constexpr (std::is_same<T, float>::value) {
        if (L.bf16) {
            build((const bf16_t*)L.coef16.data());
            return;
        }
    }
    build((const T*)L.coef.data());
}

// Calls f with the level's D^-1 in its stored type (T, or BF16).
template <class T, class F>
void with_invd(Level<T>& L, F&& f) {
    if constexpr (std::is_same<T, float>::value) {
        if (L.bf16) {
            f((const bf16_t*)L.invd16.data());
            return;
        }
    }
    f((const T*)L.invd.data());
}

// Calls f with the column-marching arguments of a hex8 level (a coefficient per cell), its
// coefficients and D^-1 in their stored type.
template <class T, class F>
void with_march(Level<T>& L, F&& f) {
    auto go = [&](auto kp, auto ip) {
        MarchArgs<T, elem_t<decltype(kp)>> m{};
        m.g = L.g;
        m.k = kp;
        m.invd = ip;
        hex_c<T>(L.h, m.c);
        f(m);
    };
    if constexpr (std::is_same<T, float>::value) {
        if (L.bf16) {
            go((const bf16_t*)L.coef16.data(), (const bf16_t*)L.invd16.data());
            return;
        }
    }
    go((const T*)L.coef.data(), (const T*)L.invd.data());
}

// The column-marching arguments of an FP64 hex8 level (the outer operator).
template <class T>
MarchArgs<T> march_args(Level<T>& L) {
    MarchArgs<T> m{};
    m.g = L.g;
    m.k = L.coef.data();
    hex_c<T>(L.h, m.c);
    return m;
}

template <class T, int MODE, class S>
void launch_march(const MarchArgs<T, S>& m, hipStream_t s) {
    if (recording_unbatchable()) return;
    hipLaunchKernelGGL((k_march_hex8<T, MODE, true, MarchShape<T>::shuffle, MarchShape<T>::kz, S>), march_grid<T>(m.g),
                       dim3(BLOCK), 0, s, m);
}

// The arguments of the fused tet10 kernels of a level (tet10_fused.hpp), with its Lame
// coefficients and D^-1 stored in S.
template <class T, class S = T>
T10Args<T, S> t10_args(const Level<T>& L) {
    using TL = T10Tile<T>;
    T10Args<T, S> a{};
    a.g = L.g;
    a.cx = L.cells[0];
    a.cy = L.cells[1];
    a.cz = L.cells[2];
    a.tx = (a.cx + TL::X - 1) / TL::X;
    a.ty = (a.cy + TL::Y - 1) / TL::Y;
    a.tz = (a.cz + TL::Z - 1) / TL::Z;
    if constexpr (sizeof(S) == 2) {
        a.coef = L.coef16.data();
        a.invd = L.invd16.data();
    } else {
        a.coef = L.coef.data();
        a.invd = L.invd.data();
    }
    a.ih0 = T(1.0 / L.h[0]);
    a.ih1 = T(1.0 / L.h[1]);
    a.ih2 = T(1.0 / L.h[2]);
    a.w = T(L.h[0] * L.h[1] * L.h[2] / 6.0 / 4.0);
    a.P = L.P.data();
    a.blist = L.blist.data();
    a.nb = L.nbound;
    return a;
}

// The partial sums of a tet10 level's fused kernels and the nodes that their boundary launch
// finishes, for the tile of T.
template <class T>
void t10_setup(Level<T>& L, hipStream_t s) {
    using TL = T10Tile<T>;
    if (8 * L.dofs() * sizeof(T) >= MARCH_MAX_BYTES || 12 * L.ncell() * sizeof(T) >= MARCH_MAX_BYTES)
        throw Error("the tet10 grid is too large for the fused operator kernels (byte offsets of 31 bits)");
    const int tn[3] = {TL::X, TL::Y, TL::Z};
    std::vector<int> list;
    t10_boundary(L.cells, tn, list);
    L.nbound = (int)list.size();
    L.blist.resize(list.size());
    if (!list.empty()) L.blist.upload(list.data(), list.size(), s);
    L.P.resize(8 * L.dofs());
    FEM_HIP(hipStreamSynchronize(s));  // the host list ends with this scope
}

// The two launches of a fused tet10 mode; returns the number of block partials (PCGQ).
template <int MODE, class T, class S>
unsigned t10_run(T10Args<T, S> a, hipStream_t s) {
    using TL = T10Tile<T>;
    const unsigned n1 = (unsigned)a.tx * a.ty * a.tz;
    const unsigned n2 = (unsigned)((a.nb + BLOCK - 1) / BLOCK);
    a.nb1 = (int)n1;
    run_k(KT10Tile<T, TL::X, TL::Y, TL::Z, MODE, S>{a}, n1, s);
    run_k(KT10Bound<T, TL::X, TL::Y, TL::Z, MODE, S>{a}, n2, s);
    return n1 + n2;
}

// Calls f with the fused tet10 arguments of a level in its storage type.
template <class T, class F>
void with_t10(const Level<T>& L, F&& f) {
    if constexpr (std::is_same<T, float>::value) {
        if (L.bf16) {
            f(t10_args<T, bf16_t>(L));
            return;
        }
    }
    f(t10_args<T>(L));
}

// y = A x on a level (0 on fixed nodes).
template <class T>
void op_apply(Level<T>& L, const T* x, T* y, hipStream_t s) {
    if (L.op == OpKind::Tet10) {
        with_t10(L, [&](auto a) {
            a.x = x;
            a.y = y;
            t10_run<T10_APPLY>(a, s);
        });
        return;
    }
    with_op(L, [&](auto op) { launch((k_apply<decltype(op), T>), L.nodes(), s, op, x, y); });
}

// r = b - A x on a level (0 on fixed nodes).
template <class T>
void op_residual(Level<T>& L, const T* x, const T* b, T* r, hipStream_t s) {
    if (L.op == OpKind::Tet10) {
        with_t10(L, [&](auto a) {
            a.x = x;
            a.b = b;
            a.y = r;
            t10_run<T10_RESID>(a, s);
        });
        return;
    }
    with_op(L, [&](auto op) { run_k(KResidual<decltype(op), T>{op, x, b, r}, nblocks(L.nodes()), s); });
}

// invd = 1 / diag(A) on a level (0 on fixed nodes).
template <class T>
void op_inv_diag(Level<T>& L, hipStream_t s) {
    if (L.op == OpKind::Tet10) {
        const T ih[3] = {T(1.0 / L.h[0]), T(1.0 / L.h[1]), T(1.0 / L.h[2])};
        const T vol = T(L.h[0] * L.h[1] * L.h[2] / 6.0);
        const VectorOp<Tet10Op<T>> op{Tet10Op<T>{L.g, L.coef.data(), {ih[0], ih[1], ih[2]}, vol}};
        launch((k_inv_diag<VectorOp<Tet10Op<T>>, T>), L.nodes(), s, op, L.invd.data());
        return;
    }
    with_op(L, [&](auto op) { launch((k_inv_diag<decltype(op), T>), L.nodes(), s, op, L.invd.data()); });
}

// Conversion of the FP64 coefficients of a level to T.
template <class T>
void convert_level(const Level<double>& src, Level<T>& dst, hipStream_t s) {
    dst.op = src.op;
    dst.comps = src.comps;
    dst.g = src.g;
    for (int d = 0; d < 3; ++d) {
        dst.cells[d] = src.cells[d];
        dst.h[d] = src.h[d];
    }
    dst.kuhn = src.kuhn;
    dst.coef.resize(src.coef.size());
    launch((k_convert<double, T>), src.coef.size(), s, src.coef.size(), src.coef.data(), dst.coef.data());
    if (src.op == OpKind::Tet10) t10_setup(dst, s);
}

// Device time of chosen launches by HIP events on the launching stream: begin/end pairs, not
// nested, category 0 the finest-level smoother step, 1 the outer FP64 operator apply. A profiled
// solve launches its iterations one by one (HIP cannot time events that graph nodes record).
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, size_t>> marks;
    double secs[2] = {0.0, 0.0};
    uint64_t counts[2] = {0, 0};

    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            FEM_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int cat, hipStream_t s) {
        if (!on) return;
        marks.push_back({cat, used});
        FEM_HIP(hipEventRecord(next(), s));
    }
    void end(hipStream_t s) {
        if (!on) return;
        FEM_HIP(hipEventRecord(next(), s));
    }
    // Adds the times of `m` (whose events have completed) to the totals.
    void flush(const std::vector<std::pair<int, size_t>>& m) {
        for (const auto& k : m) {
            FEM_HIP(hipEventSynchronize(pool[k.second + 1]));
            float ms = 0.f;
            FEM_HIP(hipEventElapsedTime(&ms, pool[k.second], pool[k.second + 1]));
            secs[k.first] += ms * 1e-3;
            counts[k.first] += 1;
        }
    }
    void flush_eager() {
        flush(marks);
        marks.clear();
        used = 0;
    }
    void reset_totals() {
        secs[0] = secs[1] = 0.0;
        counts[0] = counts[1] = 0;
    }
    ~Profiler() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

// ---------------------------------------------------------------------------------------------
// The hierarchy of one precision: levels, smoother, dense coarse level.
template <class T>
class Hierarchy {
   public:
    std::vector<Level<T>> lv;
    DenseCoarse coarse;
    bool bf16 = false;  // BF16 storage of the levels' coefficients and D^-1 (FP32 hierarchies)
    Profiler* prof = nullptr;
    // With rz_part set, level 0's last smoothing step also sums r.z of the outer PCG (r = rz_r, z
    // its output) into march_blocks partials at rz_part, when fuses_rz().
    const double* rz_r = nullptr;
    double* rz_part = nullptr;
    // Algorithmic bytes of the launches (every array a launch reads or writes, once, at its stored
    // precision), added to *ledger as they are launched.
    uint64_t* ledger = nullptr;
    void spend(uint64_t b) const {
        if (ledger) *ledger += b;
    }

    bool fuses_rz() const {
        return rz_part && lv.size() > 1 && lv[0].march() && lv[0].nodes() * sizeof(double) < MARCH_MAX_BYTES;
    }

    // The hierarchy for another solve of the same model: o's levels and coarse inverse, viewed,
    // with vectors of its own (Level::shadow_of). rz_r, rz_part and ledger are the caller's to set.
    void shadow_of(const Hierarchy& o, hipStream_t s) {
        lv.resize(o.lv.size());
        for (size_t l = 0; l < lv.size(); ++l) lv[l].shadow_of(o.lv[l], s);
        coarse.shadow_of(o.coarse, s);
        bf16 = o.bf16;
    }

    void mark_begin(size_t l, hipStream_t s) {
        if (l == 0 && prof) prof->begin(0, s);
    }
    void mark_end(size_t l, hipStream_t s) {
        if (l == 0 && prof) prof->end(s);
    }

    // first_done: the caller's last kernel has taken level 0's first Chebyshev step (x and d of
    // buffer 0, as k_cheb_first does).
    void vcycle(hipStream_t s, bool first_done = false) { cycle(0, s, first_done); }

    void cycle(size_t l, hipStream_t s, bool first_done) {
        Level<T>& L = lv[l];
        L.cur = 0;  // every cycle uses the same buffers in the same order (graph capture)
        if (l + 1 == lv.size()) {
            // first_done on the dense level: the restriction wrote its dense right side
            if (first_done) coarse.solve_rhs<T>(L.x(), s);
            else coarse.solve<T>(L.b.data(), L.x(), s);
            spend(coarse.solve_bytes(sizeof(T), !first_done));
            return;
        }
        const size_t n = L.dofs();
        const size_t vb = sizeof(T);
        // pre-smoothing from x = 0
        if (!first_done) {
            spend(3 * n * vb + L.invd_bytes());
            with_invd(L, [&](auto iv) {
                run_k(KChebFirst<T, elem_t<decltype(iv)>>{n, (const T*)L.b.data(), iv, L.x(), L.d.data(), L.c2[0]},
                      nblocks(n), s);
            });
        }
        const int deg = (int)L.c1.size();  // the level's smoothing steps
        for (int st = 1; st < deg; ++st) {
            mark_begin(l, s);
            cheb_step(L, L.c1[st], L.c2[st], s);
            mark_end(l, s);
            L.cur ^= 1;
        }
        if (L.march()) {
            with_march(L, [&](auto m) {
                m.x = L.x();
                m.b = L.b.data();
                m.y = L.r.data();
                launch_march<T, MARCH_RESIDUAL>(m, s);
            });
        } else {
            op_residual(L, (const T*)L.x(), (const T*)L.b.data(), L.r.data(), s);
        }
        spend(3 * n * vb + L.coef_bytes());
        Level<T>& Cc = lv[l + 1];
        const bool smoothed = l + 2 < lv.size();  // the next level smooths (it is not the dense one)
        transfer(L, Cc, true, s, smoothed);
        cycle(l + 1, s, true);
        transfer(L, Cc, false, s);
        for (int st = 0; st < deg; ++st) {
            // the profile times plain steps only: the fused r.z step moves 8 more bytes a node
            const bool rz = l == 0 && st + 1 == deg && fuses_rz();
            if (!rz) mark_begin(l, s);
            cheb_step(L, st == 0 ? T(0) : L.c1[st], L.c2[st], s, rz);
            if (!rz) mark_end(l, s);
            L.cur ^= 1;
        }
    }

    // d = c1 d + c2 D^-1 (b - A x), x2 = x + d; with rz (a march level), also the r.z partials.
    void cheb_step(Level<T>& L, T c1, T c2, hipStream_t s, bool rz = false) {
        // x, b, d in; d, x out; D^-1 and the coefficients; with rz the outer r (FP64) in
        spend(4 * L.dofs() * sizeof(T) + L.invd_bytes() + L.coef_bytes() + (rz ? L.dofs() * sizeof(double) : 0));
        if (L.march()) {
            with_march(L, [&](auto m) {
                m.x = L.x();
                m.y = L.x2();
                m.d = L.d.data();
                m.b = L.b.data();
                m.c1 = c1;
                m.c2 = c2;
                if (rz) {
                    m.r = rz_r;
                    m.part = rz_part;
                    launch_march<T, MARCH_CHEB_RZ>(m, s);
                } else {
                    launch_march<T, MARCH_CHEB>(m, s);
                }
            });
            return;
        }
        if (L.op == OpKind::Tet10) {
            with_t10(L, [&](auto a) {
                a.x = L.x();
                a.y = L.x2();
                a.d = L.d.data();
                a.b = L.b.data();
                a.c1 = c1;
                a.c2 = c2;
                t10_run<T10_CHEB>(a, s);
            });
            return;
        }
        with_op(L, [&](auto op) {
            with_invd(L, [&](auto iv) {
                run_k(KCheb<decltype(op), T, elem_t<decltype(iv)>>{op, (const T*)L.x(), L.x2(), L.d.data(),
                                                                 (const T*)L.b.data(), iv, c1, c2},
                      nblocks(L.nodes()), s);
            });
        });
    }

    // restrict_: b_c = P^T r_f, and with first_step the coarse level's first Chebyshev step into
    // its buffer 0 (the buffer its cycle starts from); else x_f += P x_c.
    void transfer(Level<T>& F, Level<T>& Cc, bool restrict_, hipStream_t s, bool first_step = false) {
        const int C = F.comps;
        const size_t vb = sizeof(T);
        if (restrict_) {
            const size_t n = Cc.nodes();
            // into the dense level (not smoothed): straight to its dense right side
            const bool dense = !first_step;
            // r_f in, b_c out; the coarse first step's D^-1 in and x, d out; or the dense index map in
            spend(F.dofs() * vb + Cc.dofs() * vb +
                  (first_step ? Cc.invd_bytes() + 2 * Cc.dofs() * vb : Cc.dofs() * sizeof(int32_t)));
            const int32_t* of = dense ? coarse.dense_of().data() : nullptr;
            T* db = dense ? coarse.rhs<T>() : nullptr;
            with_invd(Cc, [&](auto iv) {
                using S = elem_t<decltype(iv)>;
                const S* inv = first_step ? iv : nullptr;
                const T c0 = first_step ? Cc.c2[0] : T(0);
                T* xc = Cc.xa.data();
                T* dc = Cc.d.data();
                const T* rf = F.r.data();
                if (F.kuhn) {
                    if (C == 1) run_k(KRestrict<T, 1, true, S>{F.g, Cc.g, rf, Cc.b.data(), inv, c0, xc, dc, of, db}, nblocks(n), s);
                    else run_k(KRestrict<T, 3, true, S>{F.g, Cc.g, rf, Cc.b.data(), inv, c0, xc, dc, of, db}, nblocks(n), s);
                } else {
                    if (C == 1) run_k(KRestrict<T, 1, false, S>{F.g, Cc.g, rf, Cc.b.data(), inv, c0, xc, dc, of, db}, nblocks(n), s);
                    else run_k(KRestrict<T, 3, false, S>{F.g, Cc.g, rf, Cc.b.data(), inv, c0, xc, dc, of, db}, nblocks(n), s);
                }
            });
        } else {
            const size_t n = F.nodes();
            spend(Cc.dofs() * vb + 2 * F.dofs() * vb);  // x_c in, x_f in and out
            if (F.kuhn) {
                if (C == 1) run_k(KProlongAdd<T, 1, true>{F.g, Cc.g, (const T*)Cc.x(), F.x()}, nblocks(n), s);
                else run_k(KProlongAdd<T, 3, true>{F.g, Cc.g, (const T*)Cc.x(), F.x()}, nblocks(n), s);
            } else {
                if (C == 1) run_k(KProlongAdd<T, 1, false>{F.g, Cc.g, (const T*)Cc.x(), F.x()}, nblocks(n), s);
                else run_k(KProlongAdd<T, 3, false>{F.g, Cc.g, (const T*)Cc.x(), F.x()}, nblocks(n), s);
            }
        }
    }
};

class Batcher;

// ---------------------------------------------------------------------------------------------
class Solver {
   public:
    Solver(const BoxInput& in, const Options& o) : opt_(o), uid_(next_uid()) {
        if (in.kind == Kind::ElasticHex8)
            throw Error("hex8 elasticity is solved through its assembled matrix (fem_solver_new_box), not by the matrix-free solver");
        const auto t0 = std::chrono::steady_clock::now();
        FEM_HIP(hipSetDevice(o.device));
        hipDeviceProp_t prop;
        FEM_HIP(hipGetDeviceProperties(&prop, o.device));
        device_name_ = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ")";
        handles_from_ = solve_handles;
        handles_ = handles_from_ ? handles_from_->take() : SolveHandles::create();
        s_ = handles_.s;
        pinned_ = handles_.pinned;
        ev_ = handles_.ev;
        ev_sync_ = handles_.ev_sync;
        ev_t0_ = handles_.ev_t0;
        ev_t1_ = handles_.ev_t1;
        AllocMeter meter;
        AllocMeter* const outer_meter = alloc_meter;
        alloc_meter = &meter;
        struct Restore {
            AllocMeter* m;
            ~Restore() { alloc_meter = m; }
        } restore{outer_meter};
        use_graph_ = std::getenv("FEM_NO_GRAPH") == nullptr;  // eager launches for debugging
        model_ = build_model(in);
        const size_t nfree = model_.free.size();
        if (nfree == 0) throw Error("the model has no free unknown");
        build_levels64();
        n0_ = lv64_[0].dofs();
        x_.resize(n0_);
        b_.resize(n0_);
        r_.resize(n0_);
        z_.resize(n0_);
        p_.resize(n0_);
        q_.resize(n0_);
        tmp_.resize(n0_);
        // partials of the reducing kernels: one per block of a node, vector or march launch
        parts_ = std::max<size_t>({nblocks(n0_), nblocks(lv64_[0].nodes()), march_blocks<double>(lv64_[0].g),
                                   march_blocks<float>(lv64_[0].g)}) + 1;
        partA_.resize(parts_);
        partB_.resize(parts_);
        scal_.resize(1);
        red_.resize(4);
        free_grid_.resize(nfree);
        free_grid_.upload(model_.free_grid.data(), nfree, s_);
        fb_.resize(nfree);
        fx_.resize(nfree);
        // direct or multigrid (solve.rs:92-183): forced, or as the crossover chooses (a FLOPS
        // crossover analyses the reduced matrix's pattern in the nested dissection order)
        bool direct = false;
        if (o.method == FEM_METHOD_DIRECT) {
            direct = true;
        } else if (o.method == FEM_METHOD_AUTO) {
            if (o.crossover_kind == FEM_CROSSOVER_FLOPS) direct = auto_picks_direct(reduced_pattern(), free_xyz(), o, counts_, analysed_);
            else direct = nfree <= o.crossover;
        }
        if (direct) {
            const bool multifrontal = o.direct_solver == FEM_DIRECT_MULTIFRONTAL ||
                                      (o.direct_solver == FEM_DIRECT_AUTO && nfree > MULTIFRONTAL_FROM);
            if (multifrontal) {
                // the reduced matrix (the reference's pattern, values probed from the FP64 operator)
                // to the multifrontal factor, ordered by the free unknowns' points
                csr::Csr a;
                a.n = nfree;
                a.ptr.resize(nfree + 1);
                const uint64_t nnz = export_matrix(nullptr, nullptr, nullptr);
                a.idx.resize(nnz);
                a.val.resize(nnz);
                export_matrix(a.ptr.data(), a.idx.data(), a.val.data());
                size_t free_b = 0, total_b = 0;
                FEM_HIP(hipMemGetInfo(&free_b, &total_b));
                mf_ = std::make_unique<Multifrontal>(a, free_xyz(), s_, (uint64_t)(0.8 * (double)free_b));
            } else {
                if (nfree > DENSE_DIRECT_LIMIT) {
                    std::string why;
                    if (analysed_)
                        why = " (the crossover chose it: the factor has " + std::to_string(counts_.entries) + " entries and " +
                              rust_e(counts_.flops) + " flops; ask for multigrid)";
                    throw Error("the direct solver on the device holds up to " + std::to_string(DENSE_DIRECT_LIMIT) +
                                " unknowns; this model has " + std::to_string(nfree) + why);
                }
                direct_ = std::make_unique<DenseCoarse>();
                assemble_and_factor(lv64_[0], *direct_, false);
            }
            method_ = FEM_METHOD_DIRECT;
        } else {
            method_ = FEM_METHOD_MULTIGRID;
            if (o.precision == FEM_PREC_FP64) {
                h64_ = std::make_unique<Hierarchy<double>>();
                build_hierarchy(*h64_);
                h64_->rz_r = r_.data();
                h64_->rz_part = partA_.data();
                h64_->ledger = &ledger_;
            } else {
                h32_ = std::make_unique<Hierarchy<float>>();
                h32_->bf16 = o.precision == FEM_PREC_BF16;
                build_hierarchy(*h32_);
                h32_->rz_r = r_.data();
                h32_->rz_part = partA_.data();
                h32_->ledger = &ledger_;
                r32_.resize(n0_);
            }
        }
        FEM_HIP(hipStreamSynchronize(s_));
        setup_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        device_bytes_ = meter.peak;
    }

    // A solver of the same model and hierarchy for another right side: the model, the operator and
    // smoother data of every level and the coarse inverse are the base's (viewed: the base must
    // outlive it), its vectors, workspaces, stream and events its own. It runs exactly the base's
    // launches on its own buffers, so its result equals the base's for the same right side, bit for
    // bit. No setup is repeated.
    struct ShadowTag {};
    Solver(const Solver& base, ShadowTag) : opt_(base.opt_), uid_(next_uid()) {
        const auto t0 = std::chrono::steady_clock::now();
        if (base.method_ != FEM_METHOD_MULTIGRID) throw Error("internal: only a multigrid solver has shadows");
        FEM_HIP(hipSetDevice(opt_.device));
        handles_ = SolveHandles::create();
        s_ = handles_.s;
        pinned_ = handles_.pinned;
        ev_ = handles_.ev;
        ev_sync_ = handles_.ev_sync;
        ev_t0_ = handles_.ev_t0;
        ev_t1_ = handles_.ev_t1;
        model_ = base.model_;
        device_name_ = base.device_name_;
        method_ = base.method_;
        use_graph_ = base.use_graph_;
        max_batch_ = base.max_batch_;
        analysed_ = base.analysed_;
        counts_ = base.counts_;
        n0_ = base.n0_;
        parts_ = base.parts_;
        lv64_.resize(base.lv64_.size());
        for (size_t l = 0; l < lv64_.size(); ++l) lv64_[l].shadow_of(base.lv64_[l], s_);
        const std::pair<DArray<double>*, const DArray<double>*> vecs[] = {
            {&x_, &base.x_},         {&b_, &base.b_},         {&r_, &base.r_},   {&z_, &base.z_},
            {&p_, &base.p_},         {&q_, &base.q_},         {&tmp_, &base.tmp_}, {&partA_, &base.partA_},
            {&partB_, &base.partB_}, {&red_, &base.red_},     {&fb_, &base.fb_},   {&fx_, &base.fx_}};
        for (const auto& v : vecs) v.first->like(*v.second, s_);
        r32_.like(base.r32_, s_);
        scal_.like(base.scal_, s_);
        free_grid_.view(base.free_grid_);
        auto hook = [&](auto& H) {
            H.rz_r = r_.data();
            H.rz_part = partA_.data();
            H.ledger = &ledger_;
        };
        if (base.h32_) {
            h32_ = std::make_unique<Hierarchy<float>>();
            h32_->shadow_of(*base.h32_, s_);
            hook(*h32_);
        }
        if (base.h64_) {
            h64_ = std::make_unique<Hierarchy<double>>();
            h64_->shadow_of(*base.h64_, s_);
            hook(*h64_);
        }
        FEM_HIP(hipStreamSynchronize(s_));
        setup_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    ~Solver();

    int method() const { return method_; }
    // A number no other solver of this process has (addresses can be reused).
    uint64_t uid() const { return uid_; }
    bool crossover_counts(uint64_t* entries, double* flops) const {
        if (!analysed_) return false;
        if (entries) *entries = counts_.entries;
        if (flops) *flops = counts_.flops;
        return true;
    }

    void set_profile(bool on) {
        prof_.on = on;
        if (h32_) h32_->prof = on ? &prof_ : nullptr;
        if (h64_) h64_->prof = on ? &prof_ : nullptr;
    }

    // Device time and algorithmic bytes of the profiled launches of the last solve.
    void profile(fem_profile* p) {
        prof_.flush_eager();
        p->smoother_launches = prof_.counts[0];
        p->smoother_seconds = prof_.secs[0];
        p->outer_launches = prof_.counts[1];
        p->outer_seconds = prof_.secs[1];
        p->dofs = n0_;
        const Level<double>& L = lv64_[0];
        p->outer_bytes = 16 * n0_ + L.coef.bytes();
        if (h32_) p->smoother_bytes = smoother_bytes32(h32_->lv[0]);
        else if (h64_) p->smoother_bytes = 6 * n0_ * sizeof(double) + h64_->lv[0].coef.bytes();
        else p->smoother_bytes = 0;
    }
    // Algorithmic bytes of one FP32 Chebyshev step on level 0: x, b, d in and d, x out, D^-1 and the
    // coefficients in their stored type.
    uint64_t smoother_bytes32(const Level<float>& L) const {
        if (L.bf16) return 5 * n0_ * sizeof(float) + n0_ * sizeof(bf16_t) + L.coef16.bytes();
        return 6 * n0_ * sizeof(float) + L.coef.bytes();
    }

    const BoxModel& model() const { return model_; }
    Options& options() { return opt_; }
    const char* device_name() const { return device_name_.c_str(); }

    // Reduced right side over the free unknowns in reference order: load - A u_fixed.
    void rhs(double* out) {
        const size_t n = n0_;
        tmp_.upload(model_.fixed_value.data(), n, s_);
        // y = A u with u the fixed values: the operator's rows read the fixed nodes' values.
        apply64(tmp_.data(), r_.data());
        std::vector<double> y(n);
        r_.download(y.data(), n, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        for (size_t k = 0; k < model_.free.size(); ++k) {
            const size_t gi = model_.free_grid[k];
            out[k] = model_.load[gi] - y[gi];
        }
    }

    // y = A x over the free unknowns (reference order, host).
    void apply_host(const double* xin, double* yout) {
        const size_t nf = model_.free.size();
        fb_.upload(xin, nf, s_);
        launch(k_fill<double>, n0_, s_, n0_, tmp_.data(), 0.0);
        launch(k_scatter, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)fb_.data(), tmp_.data());
        apply64(tmp_.data(), r_.data());
        launch(k_gather, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)r_.data(), fx_.data());
        fx_.download(yout, nf, s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    // The pattern of the model's reduced matrix (values empty): every pair of free unknowns whose
    // nodes share an element, columns increasing (Csr::pattern, sparse.rs:22-65, then
    // Csr::submatrix(free), sparse.rs:108-135). fidx, if given, gets the free index of each grid
    // unknown (-1 for a fixed one).
    csr::Csr reduced_pattern(std::vector<int64_t>* fidx_out = nullptr) const {
        const BoxModel& m = model_;
        const bool quad = m.in.kind == Kind::ElasticTet10;
        const int C = m.comps;
        const int nx = m.gn[0], ny = m.gn[1], nz = m.gn[2];
        const uint64_t nf = m.free.size();
        std::vector<int64_t> fidx(m.grid_unknowns(), -1);
        for (uint64_t r = 0; r < nf; ++r) fidx[m.free_grid[r]] = (int64_t)r;
        // neighbour offsets of the pattern, per class of node (its coordinates mod 2 on tet10;
        // one class otherwise)
        std::vector<std::array<int, 3>> offs[8];
        if (quad) {
            t10_pattern_offsets(offs);
        } else {
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (m.in.kind != Kind::HeatHex8) {  // a Kuhn edge: all components >= 0 or all <= 0
                            const bool up = dx >= 0 && dy >= 0 && dz >= 0, down = dx <= 0 && dy <= 0 && dz <= 0;
                            if (!up && !down) continue;
                        }
                        offs[0].push_back({dx, dy, dz});
                    }
        }
        csr::Csr out;
        out.n = nf;
        out.ptr.assign(nf + 1, 0);
        auto row_cols = [&](uint64_t r, std::vector<uint32_t>& cols) {
            cols.clear();
            const int64_t gu = m.free_grid[r];
            const int64_t node = gu / C;
            const int i = (int)(node % nx), j = (int)((node / nx) % ny), k = (int)(node / ((int64_t)nx * ny));
            const int cls = quad ? (i & 1) + 2 * (j & 1) + 4 * (k & 1) : 0;
            for (const auto& o : offs[cls]) {
                const int a = i + o[0], b = j + o[1], c = k + o[2];
                if (a < 0 || b < 0 || c < 0 || a >= nx || b >= ny || c >= nz) continue;
                const int64_t nb = a + (int64_t)nx * (b + (int64_t)ny * c);
                for (int q = 0; q < C; ++q) {
                    const int64_t f = fidx[nb * C + q];
                    if (f >= 0) cols.push_back((uint32_t)f);
                }
            }
            std::sort(cols.begin(), cols.end());
        };
        std::vector<uint32_t> cols;
        for (uint64_t r = 0; r < nf; ++r) {
            row_cols(r, cols);
            out.ptr[r + 1] = out.ptr[r] + cols.size();
        }
        out.idx.resize(out.ptr[nf]);
        for (uint64_t r = 0; r < nf; ++r) {
            row_cols(r, cols);
            std::copy(cols.begin(), cols.end(), out.idx.begin() + out.ptr[r]);
        }
        if (fidx_out) *fidx_out = std::move(fidx);
        return out;
    }

    // The coordinates of each free unknown's node for the nested dissection of the crossover: its
    // grid indices (refined twice on tet10). The order reads coordinates only through comparisons
    // along one axis at a time (order.rs:66-139), so these give the order of the reference's
    // coordinates, which increase with the grid index along each axis.
    std::vector<std::array<double, 3>> free_xyz() const {
        const BoxModel& m = model_;
        const int nx = m.gn[0], ny = m.gn[1];
        std::vector<std::array<double, 3>> xyz(m.free.size());
        for (size_t r = 0; r < m.free.size(); ++r) {
            const int64_t node = m.free_grid[r] / m.comps;
            xyz[r] = {(double)(node % nx), (double)((node / nx) % ny), (double)(node / ((int64_t)nx * ny))};
        }
        return xyz;
    }

    // The reduced matrix of the model (Csr::submatrix(free), sparse.rs:108-135, of the matrix the
    // reference's callers assemble): its pattern is every pair of free unknowns whose nodes share
    // an element (Csr::pattern, sparse.rs:22-65: 27 neighbours on hex8 cells, the 15 of the Kuhn
    // edges on tetrahedra, and on tet10 the nodes of the tetrahedra around each node of the grid
    // refined twice), columns in increasing order; its values come from probing the FP64 operator
    // with one unit vector per colour of nodes S apart (S = 3, or 5 on tet10, whose couplings
    // reach 2 nodes). Returns the entry count; fills the arrays when they are given.
    uint64_t export_matrix(uint64_t* out_ptr, uint32_t* out_idx, double* out_val) {
        const BoxModel& m = model_;
        const bool quad = m.in.kind == Kind::ElasticTet10;
        const int C = m.comps;
        const uint64_t nf = m.free.size();
        std::vector<int64_t> fidx;
        const csr::Csr pat = reduced_pattern(&fidx);
        const std::vector<uint64_t>& ptr = pat.ptr;
        const std::vector<uint32_t>& idx = pat.idx;
        const uint64_t nnz = ptr[nf];
        if (!out_ptr) return nnz;
        DArray<uint64_t> dptr(nf + 1);
        DArray<uint32_t> didx(std::max<uint64_t>(nnz, 1));
        DArray<double> dval(std::max<uint64_t>(nnz, 1)), px(n0_), py(n0_);
        DArray<int64_t> dof(fidx.size());
        dptr.upload(ptr.data(), nf + 1, s_);
        if (nnz) didx.upload(idx.data(), nnz, s_);
        dof.upload(fidx.data(), fidx.size(), s_);
        Level<double>& L = lv64_[0];
        const int S = quad ? 5 : 3;
        for (int c0 = 0; c0 < S; ++c0)
            for (int c1 = 0; c1 < S; ++c1)
                for (int c2 = 0; c2 < S; ++c2)
                    for (int comp = 0; comp < C; ++comp) {
                        launch(k_probe_fill, L.dofs(), s_, L.g, C, S, c0, c1, c2, comp, (const int64_t*)dof.data(), px.data());
                        op_apply(L, (const double*)px.data(), py.data(), s_);
                        launch(k_probe_csr, nf, s_, nf, L.g, C, S, c0, c1, c2, comp, (const uint64_t*)dptr.data(),
                               (const uint32_t*)didx.data(), (const int64_t*)free_grid_.data(), (const double*)py.data(),
                               dval.data());
                    }
        std::copy(ptr.begin(), ptr.end(), out_ptr);
        if (out_idx) std::copy(idx.begin(), idx.end(), out_idx);
        if (out_val && nnz) dval.download(out_val, nnz, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        return nnz;
    }

    // The pattern offsets of tet10 on the Kuhn box, on the grid refined twice: the mesh repeats
    // with a period of one cell (2 nodes), so the nodes that share a tetrahedron with a node are
    // its position plus a set of offsets that depends only on its coordinates mod 2 (class
    // x + 2 y + 4 z of those parities). The sets come from the 6 tetrahedra of one cell, all pairs of
    // their 10 nodes. Near the box's faces the offsets that leave the grid are dropped; the rest
    // still share a tetrahedron inside the box, because the split is conforming across cell faces.
    static void t10_pattern_offsets(std::vector<std::array<int, 3>> offs[8]) {
        static const int ORDER[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int t = 0; t < 6; ++t) {
            int v[4][3] = {{0, 0, 0}};
            for (int s2 = 0; s2 < 3; ++s2) {
                for (int d = 0; d < 3; ++d) v[s2 + 1][d] = v[s2][d];
                v[s2 + 1][ORDER[t][s2]] += 2;
            }
            std::array<int, 3> nd[10];
            int q = 0;
            for (int a = 0; a < 4; ++a) nd[q++] = {v[a][0], v[a][1], v[a][2]};
            for (int a = 0; a < 4; ++a)
                for (int b = a + 1; b < 4; ++b) nd[q++] = {(v[a][0] + v[b][0]) / 2, (v[a][1] + v[b][1]) / 2, (v[a][2] + v[b][2]) / 2};
            for (int a = 0; a < 10; ++a)
                for (int b = 0; b < 10; ++b) {
                    const int cls = (nd[a][0] & 1) + 2 * (nd[a][1] & 1) + 4 * (nd[a][2] & 1);
                    offs[cls].push_back({nd[b][0] - nd[a][0], nd[b][1] - nd[a][1], nd[b][2] - nd[a][2]});
                }
        }
        for (int c = 0; c < 8; ++c) {
            std::sort(offs[c].begin(), offs[c].end());
            offs[c].erase(std::unique(offs[c].begin(), offs[c].end()), offs[c].end());
        }
    }

    void precondition_host(const double* rin, double* zout) {
        if (method_ != FEM_METHOD_MULTIGRID) throw Error("the direct solver has no preconditioner");
        const size_t nf = model_.free.size();
        fb_.upload(rin, nf, s_);
        launch(k_fill<double>, n0_, s_, n0_, r_.data(), 0.0);
        launch(k_scatter, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)fb_.data(), r_.data());
        precondition();
        launch(k_gather, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)z_.data(), fx_.data());
        fx_.download(zout, nf, s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    void level_info(int l, uint64_t* unknowns, double* lmax) const {
        if (h32_) {
            if (l < 0 || l >= (int)h32_->lv.size()) throw Error("no such level");
            *unknowns = free_count(h32_->lv[l].g, h32_->lv[l].comps);
            *lmax = h32_->lv[l].lmax;
        } else if (h64_) {
            if (l < 0 || l >= (int)h64_->lv.size()) throw Error("no such level");
            *unknowns = free_count(h64_->lv[l].g, h64_->lv[l].comps);
            *lmax = h64_->lv[l].lmax;
        } else {
            if (l != 0) throw Error("no such level");
            *unknowns = model_.free.size();
            *lmax = 0.0;
        }
    }

    // Solver::solve with device vectors in reference order.
    void solve_device(const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx, fem_stats* st) {
        start_run(d_b, d_x, cancelled, ctx);
        while (!advance(true)) {
        }
        if (st) *st = run_.out;
        if (!run_.error.empty()) throw Error(run_.error);
    }

    // ---- one solve as a resumable state machine: start_run, then advance until it reports done ----
    // The same launches in the same order whether advance waits (a single solve) or polls (the
    // serving loop interleaving many solves on their own streams), so a solve's result does not
    // depend on what else runs.
    void start_run(const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx) {
        const size_t nf = model_.free.size();
        run_ = Run{};
        run_.cancelled = cancelled;
        run_.ctx = ctx;
        run_.d_out = d_x;
        FEM_HIP(hipEventRecord(ev_t0_, s_));
        ledger_ = 0;
        // two fills of the grid vectors, two scatters of the free vectors (map, source, target)
        ledger_ += 2 * n0_ * sizeof(double) + 2 * nf * 3 * sizeof(double);
        launch(k_fill<double>, n0_, s_, n0_, b_.data(), 0.0);
        launch(k_fill<double>, n0_, s_, n0_, x_.data(), 0.0);
        launch(k_scatter, nf, s_, nf, (const int64_t*)free_grid_.data(), d_b, b_.data());
        launch(k_scatter, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)d_x, x_.data());
        prof_.flush_eager();
        prof_.reset_totals();
        run_.out.method = method_;
        run_.out.precision = method_ == FEM_METHOD_DIRECT ? FEM_PREC_FP64 : opt_.precision;
        run_.out.setup_seconds = setup_seconds_;
        if (method_ == FEM_METHOD_DIRECT) {
            try {
                // x = A^-1 b, then one step of refinement on the FP64 residual: x += A^-1 (b - A x)
                // (the multifrontal factor works on the free unknowns: gathered, solved, scattered)
                auto dsolve = [&](const double* rhs, double* out) {
                    if (!mf_) {
                        direct_->solve<double>(rhs, out, s_);
                        return;
                    }
                    launch(k_gather, nf, s_, nf, (const int64_t*)free_grid_.data(), rhs, fb_.data());
                    mf_->solve(fb_.data(), fx_.data(), s_);
                    launch(k_scatter, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)fx_.data(), out);
                };
                dsolve(b_.data(), x_.data());
                apply64(x_.data(), tmp_.data());
                launch(k_axpby, n0_, s_, n0_, 1.0, (const double*)b_.data(), -1.0, tmp_.data());
                launch(k_fill<double>, n0_, s_, n0_, z_.data(), 0.0);  // the solve writes the free rows only
                dsolve(tmp_.data(), z_.data());
                launch(k_axpby, n0_, s_, n0_, 1.0, (const double*)z_.data(), 1.0, x_.data());
                const double res = true_residual();
                run_.out.iterations = 0;
                run_.out.residual = res;
                run_.out.true_residual = res;
                run_.out.levels = 1;
                run_.out.coarse_unknowns = nf;
            } catch (const Error& e) {
                run_.error = e.what();
            }
            finish_run();
            return;
        }
        launch_start();
    }

    // Advances the running solve; returns true once it has finished (solution and stats out, or an
    // error). With wait = false it returns false at once while the device is still working.
    bool advance(bool wait) {
        for (;;) {
            if (run_.phase == Phase::Finished) return true;
            if (run_.want_iter) return false;  // the serving loop's next batch runs the iteration
            if (run_.pending) {
                const hipEvent_t ev = pending_ev_ ? pending_ev_ : ev_;
                if (wait) {
                    // a blocking caller polls: the batch's readback reaches the host sooner than
                    // through hipEventSynchronize (heat 128^3: 1.0-1.5 % faster solves)
                    hipError_t q;
                    while ((q = hipEventQuery(ev)) == hipErrorNotReady) {
                    }
                    FEM_HIP(q);
                } else {
                    const hipError_t q = hipEventQuery(ev);
                    if (q == hipErrorNotReady) return false;
                    FEM_HIP(q);
                }
                run_.pending = false;
                pending_ev_ = nullptr;
            }
            try {
                const PcgScalars ps = *reinterpret_cast<const PcgScalars*>(pinned_);
                run_.stopped = ps.active == 0.0;
                if (run_.phase == Phase::Start) {
                    const double rr = ps.rr, bb = ps.bb;
                    run_.norm_b = std::sqrt(bb);
                    if (run_.norm_b == 0.0) {
                        launch(k_fill<double>, n0_, s_, n0_, x_.data(), 0.0);
                        run_.out.iterations = 0;
                        run_.out.residual = 0.0;
                        run_.out.true_residual = 0.0;
                        finish_run();
                        continue;
                    }
                    const size_t n = n0_;
                    const unsigned nbv = rblocks(n);
                    PcgScalars* sc = scal_.data();
                    precondition();
                    // r.z (r, z in, partials out), the scalars, p = z (z in, p and x in and out)
                    ledger_ += 2 * n * sizeof(double) + 2 * nbv * sizeof(double) + 5 * n * sizeof(double);
                    launch_blocks((k_pcg_rz<float>), nbv, s_, n, (const float*)nullptr, z_.data(),
                                  (const double*)r_.data(), partA_.data());
                    launch_reduce(k_pcg_scalars, s_, sc, (const double*)partA_.data(), (const double*)nullptr,
                                  (int)nbv, (int)nbv, 3);
                    launch((k_pcg_dir<double>), n, s_, n, (const PcgScalars*)sc, (const double*)z_.data(), p_.data(),
                           x_.data(), 1);
                    run_.it = 0;
                    run_.residual = std::sqrt(rr) / run_.norm_b;
                    run_.rate = 0.0;
                    s_dirty_ = true;  // a batched iteration after these launches waits for them
                } else {  // Iterating: a batch of iterations has completed (or stopped on the device)
                    const uint64_t it = (uint64_t)ps.it;
                    const double res = std::sqrt(ps.rr) / run_.norm_b;
                    if (it > run_.it && run_.residual > 0.0 && res > 0.0)
                        run_.rate = std::pow(res / run_.residual, 1.0 / (double)(it - run_.it));
                    run_.it = it;
                    run_.residual = res;
                }
                next_step();
            } catch (const Error& e) {
                run_.error = e.what();
                finish_run();
            }
        }
    }

    // The device bytes of the fused tet10 kernels' partial sums and node list on a level of c
    // cells (at most every node in the list).
    template <class T>
    static uint64_t t10_record_bytes(const int c[3]) {
        const uint64_t nodes = (2 * (uint64_t)c[0] + 1) * (2 * (uint64_t)c[1] + 1) * (2 * (uint64_t)c[2] + 1);
        return nodes * 24 * sizeof(T) + nodes * 4;
    }

    // An upper estimate of the device bytes a solver of this model holds (admission by footprint in
    // the serving loop): the level-0 FP64 vectors and coefficients, every level of the V-cycle in its
    // precision, the boundary records of the fused tet10 kernels, the dense coarse inverse or direct
    // factor with its workspace, and a quarter more for partials and allocation granularity.
    static uint64_t estimate_bytes(const BoxModel& m, const Options& o) {
        const uint64_t nfree = m.free_count;
        const uint64_t cells = (uint64_t)m.in.n[0] * m.in.n[1] * m.in.n[2];
        const uint64_t nodes0 = (uint64_t)m.gn[0] * m.gn[1] * m.gn[2];
        const uint64_t dofs0 = nodes0 * m.comps;
        const uint64_t vb = o.precision == FEM_PREC_FP64 ? 8 : 4;
        uint64_t bytes = 7 * dofs0 * 8 + 3 * nfree * 8 + 2 * (dofs0 / 256 + 64) * 8;
        auto coef64 = [&](Kind k, uint64_t c) -> uint64_t {
            switch (k) {
                case Kind::HeatHex8: return c * 8;
                case Kind::HeatTet4: return 6 * c * 8;
                default: return 12 * c * 8;
            }
        };
        // the probed matrix, the factor's A, U and inverse, its panels, and the kept inverse
        auto dense = [&](uint64_t n) -> uint64_t {
            const uint64_t np = (n + 63) / 64 * 64;
            return np * np * 8 * 4 + np * 64 * 8 * 2 + np * np * vb + (1u << 18);
        };
        const bool direct = o.method == FEM_METHOD_DIRECT || (o.method == FEM_METHOD_AUTO && nfree <= o.crossover);
        bytes += coef64(m.in.kind, cells);
        if (m.in.kind == Kind::ElasticTet10) bytes += t10_record_bytes<double>(m.in.n);
        if (direct) return (bytes + dense(nfree)) * 5 / 4;
        // V-cycle levels: level 0, then (tet10) tet4 on the same cells, then halvings
        int c[3] = {m.in.n[0], m.in.n[1], m.in.n[2]};
        uint64_t dofs = dofs0;
        Kind kind = m.in.kind;
        int lev = 0;
        uint64_t last_free = nfree;
        for (;;) {
            const uint64_t lc = (uint64_t)c[0] * c[1] * c[2];
            bytes += 6 * dofs * vb;
            if (o.precision == FEM_PREC_BF16) bytes += dofs * 2 + coef64(kind, lc) / 4;
            if (lev > 0) {
                const uint64_t nodes = (uint64_t)(c[0] + 1) * (c[1] + 1) * (c[2] + 1);
                const uint64_t cf = kind == Kind::HeatHex8 && o.coarse_galerkin ? 14 * nodes * 8 + 64 * lc * 8 : coef64(kind, lc);
                bytes += cf + cf * vb / 8;
            } else {
                bytes += coef64(kind, lc) * vb / 8;
                if (kind == Kind::ElasticTet10)
                    bytes += vb == 8 ? t10_record_bytes<double>(c) : t10_record_bytes<float>(c);
            }
            ++lev;
            if (o.levels > 0 && lev >= o.levels) break;
            Kind next = kind;
            int nc[3] = {c[0], c[1], c[2]};
            if (kind == Kind::ElasticTet10) {
                next = Kind::ElasticTet4;
            } else {
                if (o.levels == 0 && last_free <= auto_coarse(nfree, o)) break;
                if (!coarsens(c)) break;
                for (int d = 0; d < 3; ++d) nc[d] = (c[d] + 1) / 2;
            }
            kind = next;
            for (int d = 0; d < 3; ++d) c[d] = nc[d];
            const uint64_t nodes = (uint64_t)(c[0] + 1) * (c[1] + 1) * (c[2] + 1);
            dofs = nodes * m.comps;
            last_free = dofs;
        }
        return (bytes + dense(std::min<uint64_t>(last_free, o.coarse_max))) * 5 / 4;
    }

    // The state of the running solve.
    bool finished() const { return run_.phase == Phase::Finished; }
    const fem_stats& run_stats() const { return run_.out; }
    const std::string& run_error() const { return run_.error; }
    // Host staging of the free-unknown vectors of a solve (the serving loop's requests).
    double* stage_b() { return fb_.data(); }
    double* stage_x() { return fx_.data(); }
    hipStream_t stream() const { return s_; }
    size_t free_count_host() const { return model_.free.size(); }
    int64_t device_bytes() const { return device_bytes_; }
    // The most PCG iterations queued between two readbacks. Solves that share the GPU keep it at
    // 1: streams share the hardware queues, and a long batch on one holds up the others.
    void set_max_batch(int k) { max_batch_ = std::max(1, k); }

    // ---- batched iterations (the serving loop) ----
    // Records the launches of one PCG iteration (none of them runs): the iteration is batchable
    // when every launch in it is a run_k one (not the hex8 marching kernels, not the direct method).
    void record_iteration() {
        if (method_ != FEM_METHOD_MULTIGRID) return;
        Recorder rec;
        const uint64_t before = ledger_;
        tl_recorder = &rec;
        try {
            iteration_body();
        } catch (...) {
            tl_recorder = nullptr;
            ledger_ = before;
            throw;
        }
        tl_recorder = nullptr;
        rec_bytes_ = ledger_ - before;
        ledger_ = before;
        if (!rec.unbatchable) recs_ = std::move(rec.recs);
    }
    bool batchable() const { return !recs_.empty(); }
    // With on, each PCG iteration waits for the serving loop to run it in a batch.
    void set_batched(bool on) { batched_ = on && batchable(); }
    bool batched() const { return batched_; }
    const std::vector<BatchRec>& iteration_records() const { return recs_; }
    bool wants_iteration() const { return run_.want_iter && run_.phase != Phase::Finished; }
    // Makes stream bs wait for the work this solve queued on its own stream since its last batched
    // iteration.
    void order_before(hipStream_t bs) {
        if (!s_dirty_) return;
        FEM_HIP(hipEventRecord(ev_sync_, s_));
        FEM_HIP(hipStreamWaitEvent(bs, ev_sync_, 0));
        s_dirty_ = false;
    }
    // The readback of the scalars on stream bs after a batched iteration.
    void read_back_on(hipStream_t bs) {
        FEM_HIP(hipMemcpyAsync(pinned_, scal_.data(), sizeof(PcgScalars), hipMemcpyDeviceToHost, bs));
    }
    // A batched iteration and its readback are queued; ev completes after them.
    void iteration_queued(hipEvent_t ev) {
        run_.want_iter = false;
        run_.pending = true;
        pending_ev_ = ev;
        ledger_ += rec_bytes_;
    }

    // k right sides at once (multi-RHS reuse of one hierarchy: transient steps, load cases):
    // b_i = d_b + i n and x_i = d_x + i n on the device (n free unknowns; x_i the start value in,
    // the solution out). This solver takes the first and k - 1 shadows (kept for later calls) the
    // others; their PCG iterations run in batches, each launch covering the same level of every
    // right side's V-cycle. Each result equals the solve of that right side alone, bit for bit.
    // A direct solver, or a hierarchy whose iteration cannot be batched (hex8 boxes), solves them
    // concurrently on their own streams instead (still from one setup). Stats per right side;
    // throws the first right side's error after all have ended.
    void solve_many_device(size_t k, const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx,
                           fem_stats* st);

    // solve_many_device on host vectors (k * n values each).
    void solve_many_host(size_t k, const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* st) {
        const size_t nf = model_.free.size();
        DArray<double> db(k * nf), dx(k * nf);
        db.upload(b, k * nf, s_);
        dx.upload(x, k * nf, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        std::string error;
        try {
            solve_many_device(k, db.data(), dx.data(), cancelled, ctx, st);
        } catch (const Error& e) {
            error = e.what();
        }
        dx.download(x, k * nf, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        if (!error.empty()) throw Error(error);
    }

    void solve_host(const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* st) {
        const size_t nf = model_.free.size();
        fb_.upload(b, nf, s_);
        fx_.upload(x, nf, s_);
        std::string error;
        try {
            solve_device(fb_.data(), fx_.data(), cancelled, ctx, st);
        } catch (const Error& e) {
            error = e.what();
        }
        fx_.download(x, nf, s_);
        FEM_HIP(hipStreamSynchronize(s_));
        if (!error.empty()) throw Error(error);
    }

   private:
    SolveHandles handles_;                // s_, pinned_ and the events below
    HandlePool* handles_from_ = nullptr;  // the pool they came from (and go back to), if any
    static uint64_t next_uid() {
        static std::atomic<uint64_t> n{0};
        return ++n;
    }
    Options opt_;
    const uint64_t uid_;
    BoxModel model_;
    hipStream_t s_ = nullptr;
    std::string device_name_;
    int method_ = FEM_METHOD_MULTIGRID;
    double setup_seconds_ = 0.0;
    std::vector<Level<double>> lv64_;  // FP64 coefficients of every level (level 0: the outer operator)
    std::unique_ptr<Hierarchy<float>> h32_;
    std::unique_ptr<Hierarchy<double>> h64_;
    std::unique_ptr<DenseCoarse> direct_;
    std::unique_ptr<Multifrontal> mf_;  // the direct method's factor above MULTIFRONTAL_FROM unknowns
    Profiler prof_;
    bool use_graph_ = true;
    int max_batch_ = 8;
    hipGraphExec_t graph_exec_ = nullptr;
    double* pinned_ = nullptr;  // host-visible scalars read back each iteration
    // Batched iterations (the serving loop): the recorded launches of one PCG iteration and their
    // algorithmic bytes; batched_ routes iterations to the loop's batches; the event of the batch
    // whose readback is pending; s_dirty_: work was queued on s_ since the last batched iteration.
    std::vector<BatchRec> recs_;
    uint64_t rec_bytes_ = 0;
    bool batched_ = false, s_dirty_ = false;
    hipEvent_t pending_ev_ = nullptr, ev_sync_ = nullptr;
    hipEvent_t ev_ = nullptr, ev_t0_ = nullptr, ev_t1_ = nullptr;  // readback; device time of a solve
    enum class Phase { Idle, Start, Iterating, Finished };
    struct Run {
        Phase phase = Phase::Idle;
        bool pending = false;  // a readback into pinned_ is in flight (ev_, or pending_ev_)
        bool want_iter = false;  // batched: the next PCG iteration waits for the serving loop's batch
        uint64_t total_it = 0, it = 0;
        int restarts = 0;
        double norm_b = 0.0, residual = 0.0;
        double rate = 0.0;     // mean residual reduction per iteration of the last batch
        bool stopped = false;  // the device-side stopping test has ended the pass
        int (*cancelled)(void*) = nullptr;
        void* ctx = nullptr;
        double* d_out = nullptr;
        fem_stats out{};
        std::string error;
    } run_;
    int64_t device_bytes_ = 0;  // peak device bytes of the construction
    bool analysed_ = false;     // a FLOPS crossover analysed the reduced matrix (counts_)
    sym::Counts counts_;
    // Algorithmic bytes of the launches of the running solve (the hierarchies add theirs), and of
    // one captured PCG iteration (each replay of its graph adds them).
    uint64_t ledger_ = 0, iter_bytes_ = 0;
    std::vector<const void*> graph_bufs_;  // graph_buffers() at the capture of graph_exec_
    size_t n0_ = 0;
    int parts_ = 0;
    DArray<double> x_, b_, r_, z_, p_, q_, tmp_, partA_, partB_, red_, fb_, fx_;
    DArray<float> r32_;
    DArray<PcgScalars> scal_;
    DArray<int64_t> free_grid_;
    // solve_many: the shadows of this solver (declared last: destroyed before what they view) and
    // the batches of their iterations
    std::unique_ptr<Batcher> batcher_;
    std::vector<std::unique_ptr<Solver>> shadows_;

    // Automatic levels stop at the first level of at most this many unknowns. The dense coarse
    // solve costs n^2 per cycle and n^3 at setup; n0 / 256 keeps its per-cycle bytes within about
    // the finest level's (measured on MI355X: a 3,645 or 5,400 unknown coarse level beats one more
    // level of a few hundred, whose launches cost about 50 us per cycle), and 1000 keeps the setup
    // of small solves short.
    static uint64_t auto_coarse(uint64_t n0, const Options& o) {
        return std::min<uint64_t>(o.coarse_max, std::max<uint64_t>(1000, n0 / 256));
    }

    // Whether a level of c cells coarsens: ceil(c / 2) cells per axis, every axis at least 2 cells.
    static bool coarsens(const int c[3]) { return c[0] >= 2 && c[1] >= 2 && c[2] >= 2; }

    static uint64_t free_count(const GridDims& g, int comps) {
        uint64_t n = 0;
        const int top[3] = {g.nx - 1, g.ny - 1, g.nz - 1};
        // nodes not on a fixed face
        for (int k = 0; k < g.nz; ++k)
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const int q[3] = {i, j, k};
                    bool fx = false;
                    for (int f = 0; f < 6; ++f)
                        if (g.fix & (1u << f)) fx = fx || (f % 2 == 0 ? q[f / 2] == 0 : q[f / 2] == top[f / 2]);
                    if (!fx) n += comps;
                }
        return n;
    }

    // ---------------------------------------------------------------- FP64 levels
    void build_levels64() {
        const BoxModel& m = model_;
        const int quad = m.quadratic ? 1 : 0;
        lv64_.clear();
        lv64_.emplace_back();
        Level<double>& L0 = lv64_[0];
        L0.comps = m.comps;
        for (int d = 0; d < 3; ++d) {
            L0.cells[d] = m.in.n[d];
            L0.h[d] = m.h[d];
        }
        L0.g = GridDims{m.gn[0], m.gn[1], m.gn[2], m.fix};
        switch (m.in.kind) {
            case Kind::HeatHex8:
                L0.op = OpKind::Hex8K;
                L0.coef.resize(m.cell_k.size());
                L0.coef.upload(m.cell_k.data(), m.cell_k.size(), s_);
                break;
            case Kind::HeatTet4:
                L0.op = OpKind::Tet4Heat;
                L0.kuhn = true;
                break;
            case Kind::ElasticTet4:
                L0.op = OpKind::Tet4Elastic;
                L0.kuhn = true;
                break;
            case Kind::ElasticTet10:
                L0.op = OpKind::Tet10;
                L0.kuhn = true;
                t10_setup(L0, s_);
                break;
            case Kind::ElasticHex8:
                throw Error("internal: hex8 elasticity has no matrix-free level");
        }
        if (m.in.kind != Kind::HeatHex8) {
            L0.coef.resize(m.tet_coef.size());
            L0.coef.upload(m.tet_coef.data(), m.tet_coef.size(), s_);
        }
        (void)quad;
    }

    // The levels below level 0 for the multigrid, in FP64: Galerkin stencils (hex8) or averaged
    // Kuhn coefficients (tets). Stops when an axis has fewer than 2 cells, at opt.levels, or when the level is
    // small enough for the dense solve and automatic levels were asked for.
    void build_coarse_levels64() {
        const BoxModel& m = model_;
        const int want = opt_.levels;
        const uint64_t n0_free = m.free_count;
        {
            const std::vector<double> P = child_interp();
            FEM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(HEX_CHILD_P), P.data(), P.size() * sizeof(double), 0,
                                           hipMemcpyHostToDevice, s_));
            int kids[6][8];
            kuhn_kids(kids);
            FEM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(KUHN_KIDS), kids, sizeof kids, 0, hipMemcpyHostToDevice, s_));
        }
        DArray<double> cellmat;  // hex8 cell matrices of the current level
        // tet10: level 1 is tet4 on the same cells and coefficients
        if (m.in.kind == Kind::ElasticTet10 && (want == 0 || want > 1)) {
            Level<double> L1;
            L1.op = OpKind::Tet4Elastic;
            L1.comps = 3;
            L1.kuhn = true;
            for (int d = 0; d < 3; ++d) {
                L1.cells[d] = m.in.n[d];
                L1.h[d] = m.h[d];
            }
            L1.g = GridDims{m.in.n[0] + 1, m.in.n[1] + 1, m.in.n[2] + 1, m.fix};
            L1.coef.resize(lv64_[0].coef.size());
            FEM_HIP(hipMemcpyAsync(L1.coef.data(), lv64_[0].coef.data(), lv64_[0].coef.bytes(), hipMemcpyDeviceToDevice, s_));
            lv64_.push_back(std::move(L1));
        }
        for (;;) {
            Level<double>& F = lv64_.back();
            const int l = (int)lv64_.size() - 1;
            if (want > 0 && l + 1 >= want) break;
            const uint64_t nfree = free_count(F.g, F.comps);
            if (want == 0 && nfree <= auto_coarse(n0_free, opt_)) break;
            // tet10 level: its grid is refined; the next level has the same cells (handled above)
            if (F.op == OpKind::Tet10) break;
            if (!coarsens(F.cells)) break;
            Level<double> Cl;
            Cl.comps = F.comps;
            Cl.kuhn = F.kuhn;
            for (int d = 0; d < 3; ++d) {
                Cl.cells[d] = (F.cells[d] + 1) / 2;
                Cl.h[d] = 2.0 * F.h[d];
            }
            const int* fc = F.cells;
            Cl.g = GridDims{Cl.cells[0] + 1, Cl.cells[1] + 1, Cl.cells[2] + 1, m.fix};
            const size_t nc = Cl.ncell();
            if (F.op == OpKind::Hex8K && !opt_.coarse_galerkin) {
                // rediscretized: the same hex8 operator with the mean conductivity of the children
                Cl.op = OpKind::Hex8K;
                Cl.coef.resize(nc);
                launch(k_hex_average, nc, s_, Cl.cells[0], Cl.cells[1], Cl.cells[2], fc[0], fc[1], fc[2],
                       (const double*)F.coef.data(), Cl.coef.data());
            } else if (F.op == OpKind::Hex8K || F.op == OpKind::Stencil27) {
                Cl.op = OpKind::Stencil27;
                DArray<double> cm(nc * 64);
                if (F.op == OpKind::Hex8K) {
                    // M_c = P_c^T K1 P_c for the fine cell matrix K1 of k = 1
                    double c[8];
                    hex_c<double>(F.h, c);
                    std::vector<double> M(8 * 64, 0.0), P = child_interp();
                    for (int ch = 0; ch < 8; ++ch)
                        for (int a = 0; a < 8; ++a)
                            for (int b = 0; b < 8; ++b) {
                                double s = 0.0;
                                for (int p = 0; p < 8; ++p)
                                    for (int q = 0; q < 8; ++q) s += P[ch * 64 + p * 8 + a] * c[p ^ q] * P[ch * 64 + q * 8 + b];
                                M[ch * 64 + a * 8 + b] = s;
                            }
                    DArray<double> dM(8 * 64);
                    dM.upload(M.data(), M.size(), s_);
                    launch(k_hex_cells_from_k, nc, s_, Cl.cells[0], Cl.cells[1], Cl.cells[2], fc[0], fc[1], fc[2],
                           (const double*)F.coef.data(), (const double*)dM.data(), cm.data());
                    FEM_HIP(hipStreamSynchronize(s_));
                } else {
                    launch(k_hex_cells_from_cells, nc, s_, Cl.cells[0], Cl.cells[1], Cl.cells[2], fc[0], fc[1], fc[2],
                           (const double*)cellmat.data(), cm.data());
                }
                Cl.coef.resize(14 * Cl.nodes());
                launch(k_stencil_from_cells, Cl.nodes(), s_, Cl.g, (const double*)cm.data(), Cl.coef.data());
                FEM_HIP(hipStreamSynchronize(s_));
                cellmat = std::move(cm);
            } else {
                Cl.op = F.op == OpKind::Tet4Heat ? OpKind::Tet4Heat : OpKind::Tet4Elastic;
                const int q = Cl.comps == 3 ? 2 : 1;
                Cl.coef.resize((size_t)q * 6 * nc);
                launch(k_kuhn_average, (size_t)q * 6 * nc, s_, Cl.cells[0], Cl.cells[1], Cl.cells[2], fc[0], fc[1], fc[2],
                       q, (const double*)F.coef.data(), Cl.coef.data());
            }
            lv64_.push_back(std::move(Cl));
        }
    }

    static std::vector<double> child_interp() {
        std::vector<double> P(8 * 64, 0.0);
        for (int ch = 0; ch < 8; ++ch)
            for (int p = 0; p < 8; ++p)
                for (int q = 0; q < 8; ++q) {
                    double w = 1.0;
                    for (int d = 0; d < 3; ++d) {
                        const int pos = ((ch >> d) & 1) + ((p >> d) & 1);  // 0, 1, 2 in fine units
                        const int qd = (q >> d) & 1;
                        const double wd = pos == 1 ? 0.5 : (pos == 0 ? (qd == 0 ? 1.0 : 0.0) : (qd == 1 ? 1.0 : 0.0));
                        w *= wd;
                    }
                    P[ch * 64 + p * 8 + q] = w;
                }
        return P;
    }

    static void kuhn_kids(int out[6][8]) {
        const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int t = 0; t < 6; ++t) {
            int m = 0;
            for (int c = 0; c < 8; ++c)
                for (int ft = 0; ft < 6; ++ft) {
                    double y[3] = {0, 0, 0};
                    int bits[3] = {0, 0, 0};
                    for (int v = 0; v < 4; ++v) {
                        if (v > 0) bits[order[ft][v - 1]] = 1;
                        for (int d = 0; d < 3; ++d) y[d] += 0.25 * bits[d];
                    }
                    double x[3];
                    for (int d = 0; d < 3; ++d) x[d] = 0.5 * (((c >> d) & 1) + y[d]);
                    const int a = order[t][0], b = order[t][1], cc = order[t][2];
                    if (x[a] > x[b] && x[b] > x[cc]) {
                        if (m >= 8) throw Error("the Kuhn refinement is not nested");
                        out[t][m++] = c * 8 + ft;
                    }
                }
            if (m != 8) throw Error("the Kuhn refinement is not nested");
        }
    }

    // ---------------------------------------------------------------- dense assembly by probing
    // dense_of: grid unknown -> dense index (free unknowns of the level in grid order), or -1.
    std::vector<int64_t> dense_index(const GridDims& g, int C, std::vector<int64_t>& free_units) {
        std::vector<int64_t> of(g.nodes() * C, -1);
        free_units.clear();
        const int top[3] = {g.nx - 1, g.ny - 1, g.nz - 1};
        for (int k = 0; k < g.nz; ++k)
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const int q[3] = {i, j, k};
                    bool fx = false;
                    for (int f = 0; f < 6; ++f)
                        if (g.fix & (1u << f)) fx = fx || (f % 2 == 0 ? q[f / 2] == 0 : q[f / 2] == top[f / 2]);
                    if (fx) continue;
                    const size_t node = i + (size_t)g.nx * (j + (size_t)g.ny * k);
                    for (int c = 0; c < C; ++c) {
                        of[node * C + c] = (int64_t)free_units.size();
                        free_units.push_back((int64_t)(node * C + c));
                    }
                }
        return of;
    }

    void assemble_and_factor(Level<double>& L, DenseCoarse& dc, bool fp32) {
        const int C = L.comps;
        std::vector<int64_t> units;
        const std::vector<int64_t> of = dense_index(L.g, C, units);
        const int64_t n = (int64_t)units.size();
        DArray<int64_t> dof(of.size());
        dof.upload(of.data(), of.size(), s_);
        DArray<double> A((size_t)n * n), px(L.dofs()), py(L.dofs());
        launch(k_fill<double>, (size_t)n * n, s_, (size_t)n * n, A.data(), 0.0);
        const int S = L.op == OpKind::Tet10 ? 5 : 3;
        for (int c0 = 0; c0 < S; ++c0)
            for (int c1 = 0; c1 < S; ++c1)
                for (int c2 = 0; c2 < S; ++c2)
                    for (int comp = 0; comp < C; ++comp) {
                        launch(k_probe_fill, L.dofs(), s_, L.g, C, S, c0, c1, c2, comp, (const int64_t*)dof.data(), px.data());
                        op_apply(L, (const double*)px.data(), py.data(), s_);
                        launch(k_probe_extract, L.dofs(), s_, L.g, C, S, c0, c1, c2, comp, (const int64_t*)dof.data(),
                               (const double*)py.data(), A.data(), n);
                    }
        FEM_HIP(hipGetLastError());
        dc.factor(A.data(), (int)n, fp32, s_);
        dc.map().resize(units.size());
        dc.map().upload(units.data(), units.size(), s_);
        std::vector<int32_t> of32(of.begin(), of.end());
        dc.dense_of().resize(of32.size());
        dc.dense_of().upload(of32.data(), of32.size(), s_);
        FEM_HIP(hipStreamSynchronize(s_));
    }

    // ---------------------------------------------------------------- hierarchy in T
    template <class T>
    void build_hierarchy(Hierarchy<T>& H) {
        build_coarse_levels64();
        if (lv64_.size() < 2 && free_count(lv64_[0].g, lv64_[0].comps) > opt_.coarse_max)
            throw Error("the multigrid has a single level with more unknowns than the dense coarse solve holds");
        Level<double>& last = lv64_.back();
        const uint64_t ncoarse = free_count(last.g, last.comps);
        if (ncoarse > opt_.coarse_max)
            throw Error("the coarsest level has " + std::to_string(ncoarse) + " unknowns, more than the dense limit " +
                        std::to_string(opt_.coarse_max) + ": ask for more levels");
        H.lv.resize(lv64_.size());
        for (size_t l = 0; l < lv64_.size(); ++l) {
            Level<T>& L = H.lv[l];
            convert_level(lv64_[l], L, s_);
            const size_t n = L.dofs();
            L.invd.resize(n);
            L.b.resize(n);
            L.xa.resize(n);
            L.xb.resize(n);
            L.d.resize(n);
            L.r.resize(n);
            for (T* v : {L.b.data(), L.xa.data(), L.xb.data(), L.d.data(), L.r.data()}) launch(k_fill<T>, n, s_, n, v, T(0));
            if (l + 1 < lv64_.size()) {
                if constexpr (std::is_same<T, float>::value) {
                    // BF16 storage: the coefficients are rounded before D^-1 and the spectrum
                    // estimate are taken, so that both see the operator the smoother applies.
                    if (H.bf16) {
                        L.bf16 = true;
                        if (L.op != OpKind::Stencil27) {
                            L.coef16.resize(L.coef.size());
                            launch(k_round_bf16, L.coef.size(), s_, L.coef.size(), L.coef.data(), L.coef16.data());
                        }
                    }
                }
                op_inv_diag(L, s_);
                if constexpr (std::is_same<T, float>::value) {
                    if (H.bf16) {
                        L.invd16.resize(n);
                        launch(k_round_bf16, n, s_, n, L.invd.data(), L.invd16.data());
                    }
                }
                L.lmax = estimate_lmax(L);
                chebyshev(L, level_degree(opt_.degree, l, lv64_[0].comps > 1));
            }
        }
        assemble_and_factor(lv64_.back(), H.coarse, sizeof(T) == 4);
    }

    template <class T>
    double estimate_lmax(Level<T>& L) {
        const size_t n = L.dofs();
        T* v = L.xa.data();
        T* w = L.xb.data();
        launch(k_random<T>, n, s_, n, v, (const T*)L.invd.data());
        const unsigned pb = std::min<unsigned>(nblocks(n), 1024);
        // The iterate is normalised on the device; the host reads |D^-1 A v| once, after the last
        // step.
        for (int it = 0; it < opt_.power_steps; ++it) {
            launch_blocks(k_dot<T>, pb, s_, n, (const T*)v, (const T*)v, partA_.data());
            launch_reduce(k_sum, s_, (const double*)partA_.data(), (int)pb, red_.data());
            launch(k_normalise<T>, n, s_, n, v, (const double*)red_.data());
            op_apply(L, (const T*)v, w, s_);
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
        const double lam = std::sqrt(nw);
        launch(k_fill<T>, n, s_, n, L.xa.data(), T(0));
        launch(k_fill<T>, n, s_, n, L.xb.data(), T(0));
        return lam;
    }

    template <class T>
    void chebyshev(Level<T>& L, int degree) {
        const double hi = opt_.cheb_upper * L.lmax, lo = opt_.cheb_lower * L.lmax;
        const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
        L.c1.assign(degree, T(0));
        L.c2.assign(degree, T(0));
        double rho = 1.0 / sigma;
        L.c1[0] = T(0);
        L.c2[0] = T(1.0 / theta);
        for (int s = 1; s < degree; ++s) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            L.c1[s] = T(rn * rho);
            L.c2[s] = T(2.0 * rn / delta);
            rho = rn;
        }
    }

    // ---------------------------------------------------------------- FP64 outer iteration
    void apply64(const double* in, double* out) {
        Level<double>& L = lv64_[0];
        op_apply(L, in, out, s_);
    }

    // z = M^-1 r (device, grid layout).
    void precondition() {
        if (h32_) {
            Level<float>& L = h32_->lv[0];
            ledger_ += 2 * n0_ * (sizeof(double) + sizeof(float));  // r to FP32 and x back to FP64
            launch((k_convert<double, float>), n0_, s_, n0_, (const double*)r_.data(), L.b.data());
            h32_->vcycle(s_);
            launch((k_convert<float, double>), n0_, s_, n0_, (const float*)L.x(), z_.data());
        } else {
            Level<double>& L = h64_->lv[0];
            ledger_ += 4 * n0_ * sizeof(double);  // the two copies
            FEM_HIP(hipMemcpyAsync(L.b.data(), r_.data(), n0_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
            h64_->vcycle(s_);
            FEM_HIP(hipMemcpyAsync(z_.data(), L.x(), n0_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        }
    }

    // r = b - A x in FP64 on level 0 with the block partials of b.b (partA_) and r.r (partB_);
    // returns their count.
    unsigned residual64(const double* x, double* r) {
        Level<double>& L = lv64_[0];
        // x, b in, r out, the coefficients (and the partials of b.b and r.r, counted by the caller)
        ledger_ += 3 * n0_ * sizeof(double) + L.coef.bytes();
        if (L.op == OpKind::Tet10) {
            op_residual(L, x, (const double*)b_.data(), r, s_);
            const unsigned pb = std::min<unsigned>(nblocks(n0_), 1024);
            launch_blocks(k_dot<double>, pb, s_, n0_, (const double*)b_.data(), (const double*)b_.data(), partA_.data());
            launch_blocks(k_dot<double>, pb, s_, n0_, (const double*)r, (const double*)r, partB_.data());
            return pb;
        }
        if (L.march()) {
            MarchArgs<double> m = march_args(L);
            m.x = x;
            m.b = b_.data();
            m.y = r;
            m.part = partA_.data();
            m.part2 = partB_.data();
            launch_march<double, MARCH_PCG_START>(m, s_);
            return (unsigned)march_blocks<double>(L.g);
        }
        const unsigned nb = nblocks(L.nodes());
        with_op(L, [&](auto op) {
            launch_blocks((k_pcg_start<decltype(op)>), nb, s_, op, x, (const double*)b_.data(), r, partA_.data(),
                          partB_.data());
        });
        return nb;
    }

    double true_residual() {
        // |b - A x| / |b| in FP64 over the free unknowns (fixed rows are 0 in both)
        const unsigned nb = residual64(x_.data(), tmp_.data());
        ledger_ += 4 * (uint64_t)nb * sizeof(double);  // the partials, written and summed
        launch_reduce(k_sum, s_, (const double*)partA_.data(), (int)nb, red_.data());
        launch_reduce(k_sum, s_, (const double*)partB_.data(), (int)nb, red_.data() + 1);
        double v[2];
        FEM_HIP(hipMemcpyAsync(v, red_.data(), 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
        FEM_HIP(hipStreamSynchronize(s_));
        const double bn = std::sqrt(v[0]), rn = std::sqrt(v[1]);
        return bn > 0.0 ? rn / bn : rn;
    }


    // One PCG iteration after the residual check (multigrid.rs:738-750): q = A p, alpha, the x and r
    // updates, z = M^-1 r, rz, beta, p = z + beta p; the next residual is in the scalars. Captured
    // once in a HIP graph and replayed (all buffers are fixed per solver); run eagerly when the
    // launches are profiled.
    // q = A p in FP64 with the block partials of p.q; returns the number of partials.
    unsigned outer_apply() {
        Level<double>& L0 = lv64_[0];
        ledger_ += 2 * n0_ * sizeof(double) + L0.coef.bytes();  // p in, q out, the FP64 coefficients
        if (L0.march()) {
            MarchArgs<double> m = march_args(L0);
            m.x = p_.data();
            m.y = q_.data();
            m.part = partA_.data();
            launch_march<double, MARCH_PCG_Q>(m, s_);
            const dim3 gm = march_grid<double>(L0.g);
            return gm.x * gm.y;
        }
        if (L0.op == OpKind::Tet10) {
            T10Args<double> a = t10_args(L0);
            a.x = p_.data();
            a.y = q_.data();
            a.part = partA_.data();
            return t10_run<T10_PCGQ>(a, s_);
        }
        with_op(L0, [&](auto op) {
            run_k(KPcgQ<decltype(op)>{op, (const double*)p_.data(), q_.data(), partA_.data()}, nblocks(L0.nodes()), s_);
        });
        return nblocks(L0.nodes());
    }

   public:
    // Mean device time of the two hot launches, each repeated `reps` times back to back on the
    // solver's stream between one pair of HIP events: the finest-level Chebyshev step of the
    // V-cycle and the outer FP64 q = A p. They run on the solver's own work vectors, whose contents
    // every solve sets afresh, so later solves are unaffected.
    void time_kernels(int reps, fem_profile* p) {
        if (reps < 1) reps = 1;
        hipEvent_t e0, e1;
        FEM_HIP(hipEventCreate(&e0));
        FEM_HIP(hipEventCreate(&e1));
        auto timed = [&](auto body) {
            body();  // warm
            FEM_HIP(hipEventRecord(e0, s_));
            for (int r = 0; r < reps; ++r) body();
            FEM_HIP(hipEventRecord(e1, s_));
            FEM_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FEM_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms * 1e-3;
        };
        fem_profile out{};
        out.dofs = n0_;
        if (h32_ && !h32_->lv[0].c1.empty()) {
            auto& H = *h32_;
            Level<float>& L = H.lv[0];
            const size_t st = L.c1.size() > 1 ? 1 : 0;
            out.smoother_seconds = timed([&] { H.cheb_step(L, L.c1[st], L.c2[st], s_); L.cur ^= 1; });
            out.smoother_bytes = smoother_bytes32(L);
        } else if (h64_ && !h64_->lv[0].c1.empty()) {
            auto& H = *h64_;
            Level<double>& L = H.lv[0];
            const size_t st = L.c1.size() > 1 ? 1 : 0;
            out.smoother_seconds = timed([&] { H.cheb_step(L, L.c1[st], L.c2[st], s_); L.cur ^= 1; });
            out.smoother_bytes = 6 * n0_ * sizeof(double) + L.coef.bytes();
        }
        out.smoother_launches = out.smoother_seconds > 0 ? (uint64_t)reps : 0;
        out.outer_seconds = timed([&] { (void)outer_apply(); });
        out.outer_launches = reps;
        out.outer_bytes = 16 * n0_ + lv64_[0].coef.bytes();
        FEM_HIP(hipEventDestroy(e0));
        FEM_HIP(hipEventDestroy(e1));
        *p = out;
    }

   private:
    template <class T>
    void update_and_vcycle(Hierarchy<T>& H) {
        const size_t n = n0_;
        Level<T>& L = H.lv[0];
        const bool smooths = H.lv.size() > 1;
        // r, q in; r out; the V-cycle's right side out (FP32); with its first step D^-1 in, x, d out
        ledger_ += 3 * n * sizeof(double) + n * sizeof(T) + (smooths ? L.invd_bytes() + 2 * n * sizeof(T) : 0) +
                   rblocks(n) * sizeof(double);
        with_invd(L, [&](auto iv) {
            using S = elem_t<decltype(iv)>;
            run_k(KPcgUpdate<T, S>{n, (const PcgScalars*)scal_.data(), r_.data(), (const double*)q_.data(), L.b.data(),
                                   partB_.data(), smooths ? iv : (const S*)nullptr, smooths ? L.c2[0] : T(0),
                                   L.xa.data(), L.d.data()},
                  rblocks(n), s_);
        });
        // r.r (partB_) waits for the scalars after r.z (what 5): the stopping test is read at the
        // next iteration's head
        H.vcycle(s_, smooths);
    }

    void iteration_body() {
        const size_t n = n0_;
        const unsigned nbv = rblocks(n);
        PcgScalars* sc = scal_.data();
        prof_.begin(1, s_);
        const unsigned nq = outer_apply();
        prof_.end(s_);
        ledger_ += 2 * (uint64_t)nq * sizeof(double);  // the p.q partials, written and summed
        run_k(KPcgScalars{sc, (const double*)partA_.data(), (const double*)nullptr, (int)nq, (int)nq, 1}, 1, s_);
        // the update, r.r and the stopping test, and the V-cycle's first step on level 0 (when
        // level 0 smooths); then the V-cycle, then r.z and beta
        // r.z: summed by level 0's last smoothing step where it fuses it, else here
        unsigned nrz = nbv;
        const size_t zb = h32_ ? sizeof(float) : sizeof(double);
        if (h32_) {
            update_and_vcycle(*h32_);
            if (h32_->fuses_rz())
                nrz = (unsigned)march_blocks<float>(h32_->lv[0].g);
            else
                run_k(KPcgRz<float>{n, (const float*)h32_->lv[0].x(), (double*)nullptr, (const double*)r_.data(),
                                    partA_.data()},
                      nbv, s_);
        } else {
            update_and_vcycle(*h64_);
            if (h64_->fuses_rz())
                nrz = (unsigned)march_blocks<double>(h64_->lv[0].g);
            else
                run_k(KPcgRz<double>{n, (const double*)h64_->lv[0].x(), (double*)nullptr, (const double*)r_.data(),
                                     partA_.data()},
                      nbv, s_);
        }
        const bool fused_rz = h32_ ? h32_->fuses_rz() : h64_->fuses_rz();
        if (!fused_rz) ledger_ += n * zb + n * sizeof(double);  // z and r in
        // the r.z and r.r partials, summed; then z in, p and x in and out
        ledger_ += ((uint64_t)nrz + nbv) * sizeof(double) + (uint64_t)nrz * sizeof(double) + n * zb + 4 * n * sizeof(double);
        run_k(KPcgScalars{sc, (const double*)partA_.data(), (const double*)partB_.data(), (int)nrz, (int)nbv, 5}, 1, s_);
        // z straight from the V-cycle's level-0 iterate (kept until the next V-cycle)
        if (h32_)
            run_k(KPcgDir<float>{n, (const PcgScalars*)sc, (const float*)h32_->lv[0].x(), p_.data(), x_.data(), 0},
                  nblocks(n), s_);
        else
            run_k(KPcgDir<double>{n, (const PcgScalars*)sc, (const double*)h64_->lv[0].x(), p_.data(), x_.data(), 0},
                  nblocks(n), s_);
    }

    // Every device allocation that a captured iteration's nodes read or write, and the pinned
    // record of the readback, by address: they are made in the constructor and freed after the
    // graph in the destructor, so the list must stay the same from capture to every replay.
    std::vector<const void*> graph_buffers() const {
        std::vector<const void*> v = {x_.data(), b_.data(), r_.data(), z_.data(), p_.data(), q_.data(), tmp_.data(),
                                      partA_.data(), partB_.data(), red_.data(), scal_.data(), r32_.data(), pinned_};
        for (const Level<double>& L : lv64_)
            for (const void* q : {(const void*)L.coef.data(), (const void*)L.P.data(), (const void*)L.blist.data()}) v.push_back(q);
        auto levels = [&](const auto& H) {
            for (const auto& L : H.lv)
                for (const void* q : {(const void*)L.coef.data(), (const void*)L.coef16.data(), (const void*)L.invd.data(),
                                      (const void*)L.invd16.data(), (const void*)L.b.data(), (const void*)L.xa.data(),
                                      (const void*)L.xb.data(), (const void*)L.d.data(), (const void*)L.r.data(),
                                      (const void*)L.P.data(), (const void*)L.blist.data()})
                    v.push_back(q);
            v.push_back(H.coarse.map().data());
            v.push_back(H.coarse.dense_of().data());
            v.push_back(H.coarse.inverse_data());
            v.push_back(H.coarse.rhs_data());
        };
        if (h32_) levels(*h32_);
        if (h64_) levels(*h64_);
        return v;
    }

    void iterate() {
        if (!use_graph_ || prof_.on) {  // HIP cannot time events recorded by graph nodes
            iteration_body();
            prof_.flush_eager();
            return;
        }
        if (graph_exec_ && graph_buffers() != graph_bufs_)
            throw Error("internal: a buffer of the captured PCG iteration was reallocated after its capture");
        if (!graph_exec_) {
            graph_bufs_ = graph_buffers();
            hipGraph_t graph = nullptr;
            const uint64_t before = ledger_;
            FEM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
            iteration_body();
            FEM_HIP(hipStreamEndCapture(s_, &graph));
            iter_bytes_ = ledger_ - before;
            FEM_HIP(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0));
            FEM_HIP(hipGraphDestroy(graph));
        } else {
            ledger_ += iter_bytes_;
        }
        FEM_HIP(hipGraphLaunch(graph_exec_, s_));
    }

    // The scalars to the host (pinned), then the event the host waits on.
    void read_back() {
        FEM_HIP(hipMemcpyAsync(pinned_, scal_.data(), sizeof(PcgScalars), hipMemcpyDeviceToHost, s_));
        FEM_HIP(hipEventRecord(ev_, s_));
        run_.pending = true;
    }

    // multigrid.rs:692-753 as steps: r = b - A x and the scalars (a pass may take the iterations
    // left of max_iter), then one readback per batch of iterations.
    void launch_start() {
        const double limit = (double)(opt_.max_iter - std::min<uint64_t>(opt_.max_iter, run_.total_it));
        hipLaunchKernelGGL(k_pcg_config, dim3(1), dim3(64), 0, s_, scal_.data(), opt_.tol, limit);
        const unsigned nbn = residual64(x_.data(), r_.data());
        ledger_ += 4 * (uint64_t)nbn * sizeof(double);  // the b.b and r.r partials, written and summed
        launch_reduce(k_pcg_scalars, s_, scal_.data(), (const double*)partA_.data(), (const double*)partB_.data(),
                      (int)nbn, (int)nbn, 0);
        read_back();
        run_.phase = Phase::Start;
    }

    // The iterations to queue before the next readback: half of those the convergence rate of the
    // last batch says are still needed. An iteration past the end changes nothing (the stopping
    // test is on the device) but costs a whole V-cycle, while a readback costs a few microseconds,
    // and the rate of MG-PCG improves as it goes, so the prediction errs long: halving it keeps the
    // queued iterations short of the end.
    int batch_size() const {
        if (max_batch_ <= 1 || !(run_.rate > 0.0 && run_.rate < 1.0) || !(run_.residual > opt_.tol)) return 1;
        const double need = std::ceil(std::log(opt_.tol / run_.residual) / std::log(run_.rate));
        const uint64_t left = opt_.max_iter - std::min<uint64_t>(opt_.max_iter, run_.total_it + run_.it);
        const double k = std::min(std::floor(0.5 * need), (double)max_batch_);
        return (int)std::max(1.0, std::min(k, (double)std::max<uint64_t>(left, 1)));
    }

    // The test of the loop head (multigrid.rs:722-737) on the current residual, then one more
    // iteration with its readback, or the end of this PCG pass.
    void next_step() {
        // The device decides when a pass ends (the same test on the same doubles); the host
        // follows it, so that a pass can never spin on iterations the device skips.
        const bool at_limit = run_.total_it + run_.it >= opt_.max_iter;
        if (run_.residual <= opt_.tol || (run_.stopped && !at_limit)) {
            run_.total_it += run_.it;
            const double tr = true_residual();
            run_.out.iterations = run_.total_it;
            run_.out.residual = run_.residual;
            run_.out.true_residual = tr;
            run_.out.restarts = run_.restarts;
            if (!opt_.refine || tr <= opt_.tol || run_.restarts >= 5) {
                finish_run();
                return;
            }
            ++run_.restarts;
            launch_start();
            return;
        }
        if (at_limit) {
            run_.out.iterations = run_.total_it + run_.it;
            run_.out.residual = run_.residual;
            run_.out.restarts = run_.restarts;
            throw Error("conjugate gradients did not converge: the residual is " + rust_e2(run_.residual) + " after " +
                        std::to_string(opt_.max_iter) + " iterations");
        }
        if (run_.cancelled && run_.cancelled(run_.ctx)) throw Error("cancelled");
        if (batched_) {
            run_.want_iter = true;
            run_.phase = Phase::Iterating;
            return;
        }
        const int k = batch_size();
        for (int q = 0; q < k; ++q) iterate();
        read_back();
        run_.phase = Phase::Iterating;
    }

    // The solution to the caller's vector, the device time, the hierarchy facts.
    void finish_run() {
        const size_t nf = model_.free.size();
        if (method_ != FEM_METHOD_DIRECT) {
            run_.out.levels = h32_ ? h32_->lv.size() : h64_->lv.size();
            run_.out.coarse_unknowns = h32_ ? h32_->coarse.n() : h64_->coarse.n();
        }
        launch(k_gather, nf, s_, nf, (const int64_t*)free_grid_.data(), (const double*)x_.data(), run_.d_out);
        ledger_ += 3 * nf * sizeof(double);
        run_.out.hbm_bytes = ledger_;
        FEM_HIP(hipEventRecord(ev_t1_, s_));
        FEM_HIP(hipEventSynchronize(ev_t1_));
        float ms = 0.f;
        FEM_HIP(hipEventElapsedTime(&ms, ev_t0_, ev_t1_));
        run_.out.seconds = ms * 1e-3;
        run_.phase = Phase::Finished;
        run_.pending = false;
    }
};



// ---------------------------------------------------------------------------------------------
// Continuous batching of PCG iterations (no counterpart in the reference). One batch runs one
// iteration of each member solve: their recorded iterations (Solver::record_iteration) are merged
// in order, so that at each step the functor type at the head of the most members goes out as one
// launch over all of them (k_run_batch), each member's launches keeping their order. One launch
// then covers the same level of every member's V-cycle, or the same PCG step. Each member's blocks
// run exactly its own launches on its own buffers, so a result is bitwise the same batched and
// alone.
//  - A new set of members is launched directly from a device table; a set that comes due again is
//    captured with its readbacks into a hipGraph and replayed (capturing costs more host time than
//    launching a set once, and sets change as often as solves start and finish).
//  - Two batches can be queued at once (slots, each with its own table and graph, on one stream):
//    the device runs one while the caller reads back and plans the other.
class Batcher {
   public:
    Batcher() {
        FEM_HIP(hipStreamCreateWithFlags(&bs_, hipStreamNonBlocking));
        FEM_HIP(hipMalloc(&dtable_, SLOTS * TABLE_BYTES));
        FEM_HIP(hipHostMalloc((void**)&htable_, SLOTS * TABLE_BYTES, hipHostMallocDefault));
        for (int k = 0; k < SLOTS; ++k) {
            Slot& sl = slots_[k];
            FEM_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
            sl.dtable = static_cast<char*>(dtable_) + k * TABLE_BYTES;
            sl.htable = htable_ + k * TABLE_BYTES;
        }
    }
    ~Batcher() {
        (void)hipStreamSynchronize(bs_);
        for (Slot& sl : slots_) {
            if (sl.graph) (void)hipGraphExecDestroy(sl.graph);
            (void)hipEventDestroy(sl.ev);
        }
        (void)hipFree(dtable_);
        (void)hipHostFree(htable_);
        (void)hipStreamDestroy(bs_);
    }
    Batcher(const Batcher&) = delete;
    Batcher& operator=(const Batcher&) = delete;

    // Marks the slots whose batch has completed as free (their members' readbacks are then on the
    // host).
    void poll() {
        for (Slot& sl : slots_) {
            if (!sl.in_flight) continue;
            const hipError_t q = hipEventQuery(sl.ev);
            if (q == hipSuccess) sl.in_flight = false;
            else if (q != hipErrorNotReady) FEM_HIP(q);
        }
    }
    bool has_free_slot() const {
        for (const Slot& sl : slots_)
            if (!sl.in_flight) return true;
        return false;
    }
    bool idle() const { return !slots_[0].in_flight && !slots_[1].in_flight; }

    // One batched iteration of each member (every member wants one) in a free slot.
    void launch(const std::vector<Solver*>& members) {
        const auto t0 = std::chrono::steady_clock::now();
        Slot* free_slot = nullptr;
        for (Slot& sl : slots_)
            if (!sl.in_flight) {
                free_slot = &sl;
                break;
            }
        if (!free_slot) throw Error("internal: a batch was launched with no free slot");
        Slot& sl = *free_slot;
        std::vector<uint64_t> ids;
        for (Solver* s : members) ids.push_back(s->uid());
        for (Solver* s : members) s->order_before(bs_);
        double graph_s = 0.0;
        if (ids == sl.members && sl.graph) {
            FEM_HIP(hipGraphLaunch(sl.graph, bs_));
        } else if (ids == sl.members) {
            const auto tg = std::chrono::steady_clock::now();
            capture(sl, members);
            graph_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tg).count();
            FEM_HIP(hipGraphLaunch(sl.graph, bs_));
        } else {
            plan(sl, members, ids);
            enqueue(sl, members);
        }
        FEM_HIP(hipEventRecord(sl.ev, bs_));
        for (Solver* s : members) s->iteration_queued(sl.ev);
        sl.in_flight = true;
        std::lock_guard<std::mutex> g(m_);
        ++st_.batches;
        st_.iterations += members.size();
        st_.launches += sl.steps.size();
        st_.solo_launches += sl.solo;
        st_.graph_seconds += graph_s;
        st_.launch_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    fem_server_batch_stats stats() {
        std::lock_guard<std::mutex> g(m_);
        return st_;
    }

   private:
    static constexpr size_t TABLE_BYTES = 4u << 20;  // the functor tables of one batch
    static constexpr int SLOTS = 2;                  // batches queued at once
    struct TickStep {  // one launch of a batch: functors and block offsets in the table
        void (*launch)(const void*, const unsigned*, unsigned, unsigned, hipStream_t);
        size_t off_k, off_s;
        unsigned n, blocks;
    };
    struct Slot {
        hipEvent_t ev = nullptr;         // the end of its last batch
        char* dtable = nullptr;          // its table (device) and host staging
        uint8_t* htable = nullptr;
        hipGraphExec_t graph = nullptr;  // its planned batch, once captured
        std::vector<uint64_t> members;   // the solvers (uid) of the planned batch
        std::vector<TickStep> steps;
        size_t solo = 0;                 // their launches one solve at a time
        bool in_flight = false;
    };

    // The batch's launches (from the device table) and the members' readbacks, on bs_.
    void enqueue(const Slot& sl, const std::vector<Solver*>& members) {
        for (const TickStep& st : sl.steps)
            st.launch(sl.dtable + st.off_k, reinterpret_cast<const unsigned*>(sl.dtable + st.off_s), st.n, st.blocks, bs_);
        for (Solver* s : members) s->read_back_on(bs_);
    }

    void capture(Slot& sl, const std::vector<Solver*>& members) {
        hipGraph_t graph = nullptr;
        FEM_HIP(hipStreamBeginCapture(bs_, hipStreamCaptureModeThreadLocal));
        enqueue(sl, members);
        FEM_HIP(hipStreamEndCapture(bs_, &graph));
        FEM_HIP(hipGraphInstantiate(&sl.graph, graph, nullptr, nullptr, 0));
        FEM_HIP(hipGraphDestroy(graph));
        std::lock_guard<std::mutex> g(m_);
        ++st_.graphs;
    }

    // The merged launches of a new set of members and their device table.
    void plan(Slot& sl, const std::vector<Solver*>& members, const std::vector<uint64_t>& ids) {
        const size_t m = members.size();
        std::vector<size_t> head(m, 0);
        std::vector<TickStep> steps;
        size_t bytes = 0, solo = 0;
        auto align = [&](size_t a) { bytes = (bytes + a - 1) / a * a; };
        for (Solver* s : members) solo += s->iteration_records().size();
        for (;;) {
            // the functor type at the head of the most members (the first such in member order)
            decltype(TickStep::launch) best = nullptr;
            size_t best_count = 0;
            for (size_t i = 0; i < m; ++i) {
                const auto& ri = members[i]->iteration_records();
                if (head[i] >= ri.size()) continue;
                const auto f = ri[head[i]].launch;
                size_t c = 0;
                for (size_t j = 0; j < m; ++j) {
                    const auto& rj = members[j]->iteration_records();
                    if (head[j] < rj.size() && rj[head[j]].launch == f) ++c;
                }
                if (c > best_count) {
                    best = f;
                    best_count = c;
                }
            }
            if (!best) break;
            TickStep st{best, 0, 0, 0, 0};
            align(64);
            st.off_k = bytes;
            std::vector<unsigned> starts(1, 0);
            for (size_t i = 0; i < m; ++i) {
                const auto& ri = members[i]->iteration_records();
                if (head[i] >= ri.size() || ri[head[i]].launch != best) continue;
                const BatchRec& rec = ri[head[i]];
                if (bytes + rec.size > TABLE_BYTES) throw Error("internal: the batch table outgrew its buffer");
                std::memcpy(sl.htable + bytes, rec.bytes.data(), rec.size);
                bytes += rec.size;
                starts.push_back(starts.back() + rec.blocks);
                ++head[i];
                ++st.n;
            }
            align(4);
            st.off_s = bytes;
            if (bytes + starts.size() * sizeof(unsigned) > TABLE_BYTES)
                throw Error("internal: the batch table outgrew its buffer");
            std::memcpy(sl.htable + bytes, starts.data(), starts.size() * sizeof(unsigned));
            bytes += starts.size() * sizeof(unsigned);
            st.blocks = starts.back();
            steps.push_back(st);
        }
        if (sl.graph) {
            FEM_HIP(hipGraphExecDestroy(sl.graph));
            sl.graph = nullptr;
        }
        // the slot's previous batch has completed, so its table is free
        FEM_HIP(hipMemcpyAsync(sl.dtable, sl.htable, bytes, hipMemcpyHostToDevice, bs_));
        sl.members = ids;
        sl.steps = std::move(steps);
        sl.solo = solo;
    }

    hipStream_t bs_ = nullptr;  // the batches' stream
    void* dtable_ = nullptr;    // the slots' functor tables (device) and their host staging
    uint8_t* htable_ = nullptr;
    Slot slots_[SLOTS];
    std::mutex m_;  // st_
    fem_server_batch_stats st_{};
};

Solver::~Solver() {
    if (s_) (void)hipStreamSynchronize(s_);  // nothing of this solver may still run on the device
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (handles_from_) handles_from_->give(handles_);
    else handles_.destroy();
}

void Solver::solve_many_device(size_t k, const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx,
                           fem_stats* st) {
    const size_t nf = model_.free.size();
    if (k == 0) return;
    if (method_ != FEM_METHOD_MULTIGRID || k == 1) {
        std::string first;
        for (size_t i = 0; i < k; ++i) {
            try {
                solve_device(d_b + i * nf, d_x + i * nf, cancelled, ctx, st ? st + i : nullptr);
            } catch (const Error& e) {
                if (st) st[i] = run_.out;
                if (first.empty()) first = e.what();
            }
        }
        if (!first.empty()) throw Error(first);
        return;
    }
    FEM_HIP(hipSetDevice(opt_.device));
    while (shadows_.size() + 1 < k) shadows_.push_back(std::make_unique<Solver>(*this, ShadowTag{}));
    std::vector<Solver*> all{this};
    for (size_t i = 0; i + 1 < k; ++i) all.push_back(shadows_[i].get());
    bool batch = true;
    for (Solver* p : all) {
        if (!p->batchable()) p->record_iteration();
        batch = batch && p->batchable();
    }
    if (batch && !batcher_) batcher_ = std::make_unique<Batcher>();
    struct Unbatch {  // the solvers run alone again after this call, whatever happens
        std::vector<Solver*>& all;
        ~Unbatch() {
            for (Solver* p : all) p->set_batched(false);
        }
    } unbatch{all};
    for (size_t i = 0; i < k; ++i) {
        all[i]->set_batched(batch);
        all[i]->start_run(d_b + i * nf, d_x + i * nf, cancelled, ctx);
    }
    for (;;) {
        if (batch) batcher_->poll();
        bool done = true;
        for (Solver* p : all) {
            if (!p->finished()) p->advance(false);
            done = done && p->finished();
        }
        if (done) break;
        if (!batch) continue;
        while (batcher_->has_free_slot()) {
            std::vector<Solver*> members;
            for (Solver* p : all)
                if (p->wants_iteration()) members.push_back(p);
            if (members.empty()) break;
            batcher_->launch(members);
        }
    }
    std::string first;
    for (size_t i = 0; i < k; ++i) {
        if (st) st[i] = all[i]->run_.out;
        if (first.empty() && !all[i]->run_.error.empty()) first = all[i]->run_.error;
    }
    if (!first.empty()) throw Error(first);
}

// ---------------------------------------------------------------------------------------------
// The serving loop: a persistent request loop over Solver runs (no counterpart in the reference;
// its analogues are Solver reuse across right sides, heat.rs:697-735, the cancel closure of
// multigrid.rs:735 and refusal by size before allocation, heat.rs:1180-1186).
//
//  - submit() queues a solve with its model, options and right side (copied).
//  - Admission in arrival order by device footprint: a solve is admitted while the admitted ones
//    hold less than the budget (estimated bytes) and fewer than max_active are on the device.
//  - Setup threads build each admitted solve (model, hierarchy, factors: host work and device
//    allocation on the solve's own stream) and starts its run; the loop thread then advances every
//    running solve without waiting: each advance launches one PCG iteration (a replayed hipGraph)
//    on that solve's stream and returns, so the iterations of all running solves are in flight on
//    the device together, and a solve that finishes frees its slot and memory for the next one.
//  - Each run is the same sequence of launches as a solve on its own, on its own stream and
//    buffers, so its result is bitwise the same alone and in any mix.
struct ServeRequest {
    uint64_t id = 0;
    BoxInput in;
    Options opt;
    std::vector<double> b;  // empty: the model's own right side
    std::atomic<bool> cancel{false};
    std::unique_ptr<Solver> solver;
    std::vector<double> x;
    fem_stats st{};
    std::string err;          // written only under the server's lock
    std::string setup_err;    // the setup thread's failure, handed to the loop with the request
    bool done = false;
    uint64_t footprint = 0;
    int64_t device_bytes = 0;
    double t_submit = 0, t_admit = 0, t_start = 0, t_finish = 0;
};

class Server {
   public:
    // The device memory of every admitted solve (its vectors, hierarchy, coarse factor and the
    // factor's temporaries) comes from one arena of `budget` bytes reserved here.
    Server(uint64_t budget, int max_active, int device, int setup_threads, bool batch)
        : budget_(budget), max_active_(max_active > 0 ? max_active : 1), device_(device), batch_(batch),
          t0_(clock::now()) {
        FEM_HIP(hipSetDevice(device_));
        arena_ = std::make_unique<DeviceArena>(budget_);
        batcher_ = std::make_unique<Batcher>();
        loop_ = std::thread([this] { loop(); });
        reaper_ = std::thread([this] { reap_loop(); });
        for (int k = 0; k < std::max(1, setup_threads); ++k) setup_.emplace_back([this] { setup_loop(); });
    }
    // Stops the loop and the setup threads; a solve that had not finished ends with an error, and
    // the destructor returns once no thread is still inside wait().
    ~Server() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        setup_cv_.notify_all();
        loop_.join();
        reap_cv_.notify_all();
        reaper_.join();  // after the loop: it drains the requests the loop handed over
        for (auto& t : setup_) t.join();
        batcher_.reset();
        std::unique_lock<std::mutex> g(m_);
        for (auto& kv : all_)
            if (!kv.second->done) {
                kv.second->err = "the server stopped before the solve finished";
                kv.second->done = true;
            }
        done_cv_.notify_all();
        quiet_cv_.wait(g, [&] { return waiters_ == 0; });
    }

    uint64_t arena_bytes() const { return arena_->size(); }
    void batch_stats(fem_server_batch_stats* o) {
        *o = batcher_->stats();
        std::lock_guard<std::mutex> g(m_);
        o->finish_seconds = finish_seconds_;
    }
    uint64_t arena_peak() { return arena_->peak(); }
    uint64_t arena_misses() { return arena_->misses(); }

    uint64_t submit(const BoxInput& in, const Options& o, const double* b) {
        auto r = std::make_shared<ServeRequest>();
        r->in = in;
        r->opt = o;
        r->opt.device = device_;
        const BoxModel m = box_shape(in);
        if (b) r->b.assign(b, b + m.free_count);
        r->footprint = Solver::estimate_bytes(m, r->opt);
        if (r->footprint > budget_)
            throw Error("the solve needs about " + std::to_string(r->footprint) + " device bytes, more than the server's budget of " +
                        std::to_string(budget_));
        std::lock_guard<std::mutex> g(m_);
        if (!fatal_.empty()) throw Error(fatal_);
        r->id = ++next_id_;
        r->t_submit = now();
        all_[r->id] = r;
        queue_.push_back(r);
        cv_.notify_all();
        return r->id;
    }

    // 0 done (results out), 1 the solve failed (message out), 2 timeout; the request is forgotten
    // once its result has been taken.
    int wait(uint64_t id, double timeout_s, double* x, uint64_t n, fem_stats* st, fem_request_info* info,
             std::string& err) {
        std::unique_lock<std::mutex> g(m_);
        ++waiters_;
        struct Leave {
            Server* s;
            ~Leave() {
                if (--s->waiters_ == 0) s->quiet_cv_.notify_all();
            }
        } leave{this};
        auto it = all_.find(id);
        if (it == all_.end()) {
            err = "no such request";
            return 1;
        }
        std::shared_ptr<ServeRequest> r = it->second;
        auto ready = [&] { return r->done; };
        if (timeout_s < 0) {
            done_cv_.wait(g, ready);
        } else if (!done_cv_.wait_for(g, std::chrono::duration<double>(timeout_s), ready)) {
            return 2;
        }
        all_.erase(id);
        if (st) *st = r->st;
        if (info) {
            info->submit_time = r->t_submit;
            info->admit_time = r->t_admit;
            info->start_time = r->t_start;
            info->finish_time = r->t_finish;
            info->footprint = r->footprint;
            info->device_bytes = r->device_bytes;
        }
        if (x && n == r->x.size()) std::copy(r->x.begin(), r->x.end(), x);
        if (!r->err.empty()) {
            err = r->err;
            return 1;
        }
        if (x && n != r->x.size()) {
            err = "x needs " + std::to_string(r->x.size()) + " values";
            return 1;
        }
        return 0;
    }

    bool cancel(uint64_t id) {
        std::lock_guard<std::mutex> g(m_);
        auto it = all_.find(id);
        if (it == all_.end()) return false;
        it->second->cancel = true;
        return true;
    }

   private:
    using clock = std::chrono::steady_clock;
    double now() const { return std::chrono::duration<double>(clock::now() - t0_).count(); }

    static int cancel_fn(void* ctx) { return static_cast<ServeRequest*>(ctx)->cancel.load() ? 1 : 0; }

    // Admission (under m_): FIFO while the footprint fits and a slot is free.
    void admit() {
        while (!queue_.empty() && admitted_ < max_active_ && used_ + queue_.front()->footprint <= budget_) {
            auto r = queue_.front();
            queue_.pop_front();
            used_ += r->footprint;
            ++admitted_;
            r->t_admit = now();
            to_setup_.push_back(r);
        }
        if (!to_setup_.empty()) setup_cv_.notify_one();
    }

    // The end of a request: the loop thread hands it to the reaper thread, which takes its result
    // off the device and releases the solver (a copy, a stream synchronization and frees that can
    // wait for the device), so that the loop thread goes on launching batches meanwhile.
    void complete(const std::shared_ptr<ServeRequest>& r, std::string err) {
        {
            std::lock_guard<std::mutex> g(m_);
            reap_.emplace_back(r, std::move(err));
        }
        reap_cv_.notify_one();
    }

    void reap_loop() {
        try {
            FEM_HIP(hipSetDevice(device_));
            for (;;) {
                std::pair<std::shared_ptr<ServeRequest>, std::string> job;
                {
                    std::unique_lock<std::mutex> g(m_);
                    reap_cv_.wait(g, [&] { return !reap_.empty() || (stop_ && loop_done_); });
                    if (reap_.empty()) return;
                    job = std::move(reap_.front());
                    reap_.pop_front();
                }
                finish(job.first, std::move(job.second));
            }
        } catch (const std::exception& e) {
            fail_all(e.what());
        }
    }

    // The end of a request on the reaper thread: its result leaves the device before the solver
    // (whose memory goes back to the arena) does, then it is published under the lock, unless a
    // failure of the whole server has already ended it; then its slot and footprint are free.
    void finish(const std::shared_ptr<ServeRequest>& r, std::string err) {
        const auto t0 = clock::now();
        fem_stats st{};
        std::vector<double> x;
        if (err.empty()) err = r->setup_err;
        if (r->solver) {
            st = r->solver->run_stats();
            if (err.empty()) err = r->solver->run_error();
            const size_t nf = r->solver->free_count_host();
            x.resize(nf);
            FEM_HIP(hipMemcpyAsync(x.data(), r->solver->stage_x(), nf * sizeof(double), hipMemcpyDeviceToHost,
                                   r->solver->stream()));
            FEM_HIP(hipStreamSynchronize(r->solver->stream()));
            r->solver.reset();
        }
        std::lock_guard<std::mutex> g(m_);
        finish_seconds_ += std::chrono::duration<double>(clock::now() - t0).count();
        r->t_finish = now();
        if (!r->done) {
            r->st = st;
            r->x = std::move(x);
            r->err = err;
            r->done = true;
        }
        used_ -= r->footprint;
        --admitted_;
        admit();
        done_cv_.notify_all();
    }

    // A failure outside any one solve (the device itself) ends the server: every waiter gets it.
    void fail_all(const std::string& what) {
        std::lock_guard<std::mutex> g(m_);
        fatal_ = what;
        for (auto& kv : all_) {
            if (!kv.second->done) {
                kv.second->err = what;
                kv.second->done = true;
            }
        }
        stop_ = true;
        cv_.notify_all();
        setup_cv_.notify_all();
        done_cv_.notify_all();
    }

    void setup_loop() {
        try {
            setup_body();
        } catch (const std::exception& e) {
            fail_all(e.what());
        }
    }

    void setup_body() {
        FEM_HIP(hipSetDevice(device_));
        alloc_arena = arena_.get();  // every solver this thread builds takes its memory from the arena
        solve_handles = &handles_;   // and its stream, readback record and events from the pool
        for (;;) {
            std::shared_ptr<ServeRequest> r;
            {
                std::unique_lock<std::mutex> g(m_);
                setup_cv_.wait(g, [&] { return stop_ || !to_setup_.empty(); });
                if (stop_) return;
                r = to_setup_.front();
                to_setup_.pop_front();
            }
            try {
                r->solver = std::make_unique<Solver>(r->in, r->opt);
                r->device_bytes = r->solver->device_bytes();
                Solver& s = *r->solver;
                s.set_max_batch(1);
                if (batch_) {
                    s.record_iteration();
                    s.set_batched(true);
                }
                const size_t nf = s.free_count_host();
                if (r->b.empty()) {
                    r->b.resize(nf);
                    s.rhs(r->b.data());
                }
                std::vector<double> zero(nf, 0.0);
                FEM_HIP(hipMemcpyAsync(s.stage_b(), r->b.data(), nf * sizeof(double), hipMemcpyHostToDevice, s.stream()));
                FEM_HIP(hipMemcpyAsync(s.stage_x(), zero.data(), nf * sizeof(double), hipMemcpyHostToDevice, s.stream()));
                FEM_HIP(hipStreamSynchronize(s.stream()));
                r->t_start = now();
                s.start_run(s.stage_b(), s.stage_x(), cancel_fn, r.get());
            } catch (const std::exception& e) {
                r->setup_err = e.what();
                r->solver.reset();
            }
            std::lock_guard<std::mutex> g(m_);
            ready_.push_back(r);
            cv_.notify_all();
        }
    }

    void loop() {
        try {
            loop_body();
        } catch (const std::exception& e) {
            fail_all(e.what());
        }
        std::lock_guard<std::mutex> g(m_);
        loop_done_ = true;
        reap_cv_.notify_all();
    }

    void loop_body() {
        FEM_HIP(hipSetDevice(device_));
        std::vector<std::shared_ptr<ServeRequest>> active;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                if (stop_) return;
                admit();
                for (auto& r : ready_) active.push_back(r);
                ready_.clear();
                if (active.empty()) {
                    cv_.wait(g, [&] {
                        return stop_ || !ready_.empty() ||
                               (!queue_.empty() && admitted_ < max_active_ && used_ + queue_.front()->footprint <= budget_);
                    });
                    continue;
                }
            }
            batcher_->poll();
            bool progressed = false;
            for (size_t k = 0; k < active.size();) {
                auto& r = active[k];
                bool finished = !r->solver;  // setup failed
                std::string err;
                if (!finished) {
                    try {
                        finished = r->solver->advance(false);
                    } catch (const std::exception& e) {
                        err = e.what();
                        finished = true;
                    }
                }
                if (finished) {
                    complete(r, err);
                    active[k] = active.back();
                    active.pop_back();
                    progressed = true;
                } else {
                    ++k;
                }
            }
            while (batcher_->has_free_slot()) {
                std::vector<Solver*> members;
                for (auto& r : active)
                    if (r->solver && r->solver->wants_iteration()) members.push_back(r->solver.get());
                if (members.empty()) break;
                batcher_->launch(members);
                progressed = true;
            }
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }

    const uint64_t budget_;
    const int max_active_;
    const int device_;
    const bool batch_;
    std::unique_ptr<Batcher> batcher_;  // the batches of the running solves' iterations
    double finish_seconds_ = 0.0;       // the reaper's time finishing solves
    const clock::time_point t0_;
    std::unique_ptr<DeviceArena> arena_;
    HandlePool handles_;  // declared before the requests: outlives every solver
    std::mutex m_;
    std::condition_variable cv_, setup_cv_, done_cv_, quiet_cv_, reap_cv_;
    std::deque<std::pair<std::shared_ptr<ServeRequest>, std::string>> reap_;  // finished, not yet released
    bool loop_done_ = false;
    std::thread reaper_;
    int waiters_ = 0;  // threads inside wait()
    bool stop_ = false;
    std::string fatal_;
    uint64_t next_id_ = 0, used_ = 0;
    int admitted_ = 0;
    std::deque<std::shared_ptr<ServeRequest>> queue_, to_setup_;
    std::vector<std::shared_ptr<ServeRequest>> ready_;
    std::map<uint64_t, std::shared_ptr<ServeRequest>> all_;
    std::thread loop_;
    std::vector<std::thread> setup_;
};
}  // namespace fem

// ---------------------------------------------------------------------------------------------
// C ABI
// A box solver (impl) or an assembled one (csr).
struct fem_solver {
    std::unique_ptr<fem::Solver> impl;
    std::unique_ptr<fem::CsrSolver> csr;
    // A box model solved through its assembled matrix (hex8 elasticity): the model, its right
    // side and its reduced matrix, beside `csr`.
    std::unique_ptr<fem::BoxModel> box;
    std::vector<double> box_rhs;
    fem::csr::Csr box_matrix;
};

namespace {
void put_err(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}

fem::BoxInput box_input(const fem_box* box);

// The solver options of a fem_options.
fem::Options options_from(const fem_options& opt) {
    fem::Options o;
    o.tol = opt.tol;
    o.max_iter = opt.max_iter;
    o.method = opt.method;
    o.precision = opt.precision;
    o.crossover = opt.crossover;
    o.levels = opt.levels;
    o.degree = opt.cheb_degree > 0 ? opt.cheb_degree : 0;
    o.cheb_lower = opt.cheb_lower;
    o.cheb_upper = opt.cheb_upper;
    o.power_steps = opt.power_steps > 0 ? opt.power_steps : 20;
    o.coarse_max = opt.coarse_max;
    o.refine = opt.refine;
    o.device = opt.device;
    o.coarse_galerkin = opt.coarse_galerkin;
    o.crossover_kind = opt.crossover_kind;
    o.crossover_flops = opt.crossover_flops;
    o.crossover_entries = opt.crossover_entries;
    o.crossover_flat = opt.crossover_flat != 0;
    o.direct_solver = opt.direct_solver;
    if (o.direct_solver < FEM_DIRECT_AUTO || o.direct_solver > FEM_DIRECT_MULTIFRONTAL) throw fem::Error("unknown direct solver");
    if (o.precision < 0 || o.precision > 2) throw fem::Error("unknown precision");
    if (o.crossover_kind != FEM_CROSSOVER_UNKNOWNS && o.crossover_kind != FEM_CROSSOVER_FLOPS)
        throw fem::Error("unknown crossover");
    return o;
}

template <class F>
int guarded(char* err, size_t errlen, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return 1;
    }
}
}  // namespace

extern "C" {

void fem_options_default(fem_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    fem::Options d;
    o->tol = d.tol;
    o->max_iter = d.max_iter;
    o->method = d.method;
    o->precision = d.precision;
    o->crossover = d.crossover;
    o->levels = d.levels;
    o->cheb_degree = d.degree;
    o->cheb_lower = d.cheb_lower;
    o->cheb_upper = d.cheb_upper;
    o->power_steps = d.power_steps;
    o->coarse_max = d.coarse_max;
    o->refine = d.refine;
    o->device = d.device;
    o->coarse_galerkin = d.coarse_galerkin;
    o->crossover_kind = d.crossover_kind;
    o->crossover_flops = d.crossover_flops;
    o->crossover_entries = d.crossover_entries;
    o->crossover_flat = d.crossover_flat ? 1 : 0;
    o->direct_solver = d.direct_solver;
}

int fem_solver_new_box(const fem_box* box, const fem_options* opt, fem_solver** out, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!out) throw fem::Error("an output handle is needed");
        *out = nullptr;
        const fem::BoxInput in = box_input(box);
        fem_options def;
        fem_options_default(&def);
        const fem::Options o = options_from(opt ? *opt : def);
        auto* s = new fem_solver;
        try {
            if (in.kind == fem::Kind::ElasticHex8) {
                // elastic.rs:287-547: the assembled matrix, the grid level and the cells to Solver::new
                auto m = std::make_unique<fem::BoxModel>(fem::build_model(in));
                fem::HexBoxSystem sys = fem::hex_box_system(*m);
                fem::CsrUnknowns un;
                un.kind = FEM_UNKNOWNS_GRID;
                un.grid = std::move(sys.level);
                un.cells = true;
                un.cell_data = std::move(sys.cells);
                s->box_matrix = sys.a;
                s->csr = std::make_unique<fem::CsrSolver>(std::move(sys.a), std::move(un), o);
                s->box_rhs = std::move(sys.rhs);
                s->box = std::move(m);
            } else {
                s->impl = std::make_unique<fem::Solver>(in, o);
            }
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fem_solver_new_csr(const fem_csr* a, const fem_unknowns* u, const fem_options* opt, fem_solver** out, char* err,
                       size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!out) throw fem::Error("an output handle is needed");
        *out = nullptr;
        fem::csr::Csr m = fem::csr_from(a);
        fem::CsrUnknowns un = fem::unknowns_from(u, m.n);
        fem_options def;
        fem_options_default(&def);
        const fem::Options o = options_from(opt ? *opt : def);
        auto* s = new fem_solver;
        try {
            s->csr = std::make_unique<fem::CsrSolver>(std::move(m), std::move(un), o);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fem_solver_solve(fem_solver* s, const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* stats,
                     char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) s->csr->solve_host(b, x, cancelled, ctx, stats);
        else s->impl->solve_host(b, x, cancelled, ctx, stats);
    });
}

int fem_solver_solve_device(fem_solver* s, const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx,
                            fem_stats* stats, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) s->csr->solve_device(d_b, d_x, cancelled, ctx, stats);
        else s->impl->solve_device(d_b, d_x, cancelled, ctx, stats);
    });
}

int fem_solver_solve_many(fem_solver* s, uint64_t k, const double* b, double* x, int (*cancelled)(void*), void* ctx,
                          fem_stats* stats, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) {
            const size_t n = s->csr->n();
            std::string first;
            for (uint64_t i = 0; i < k; ++i) {
                try {
                    s->csr->solve_host(b + i * n, x + i * n, cancelled, ctx, stats ? stats + i : nullptr);
                } catch (const fem::Error& e) {
                    if (first.empty()) first = e.what();
                }
            }
            if (!first.empty()) throw fem::Error(first);
            return;
        }
        s->impl->solve_many_host(k, b, x, cancelled, ctx, stats);
    });
}

int fem_solver_solve_many_device(fem_solver* s, uint64_t k, const double* d_b, double* d_x, int (*cancelled)(void*),
                                 void* ctx, fem_stats* stats, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) {
            const size_t n = s->csr->n();
            std::string first;
            for (uint64_t i = 0; i < k; ++i) {
                try {
                    s->csr->solve_device(d_b + i * n, d_x + i * n, cancelled, ctx, stats ? stats + i : nullptr);
                } catch (const fem::Error& e) {
                    if (first.empty()) first = e.what();
                }
            }
            if (!first.empty()) throw fem::Error(first);
            return;
        }
        s->impl->solve_many_device(k, d_b, d_x, cancelled, ctx, stats);
    });
}

int fem_solver_crossover_counts(const fem_solver* s, uint64_t* entries, double* flops) {
    if (!s) return 1;
    const bool ok = s->impl ? s->impl->crossover_counts(entries, flops) : s->csr->crossover_counts(entries, flops);
    return ok ? 0 : 1;
}

int fem_solver_method(const fem_solver* s) {
    if (!s) return -1;
    return s->csr ? s->csr->method() : s->impl->method();
}

void fem_solver_set_max_iter(fem_solver* s, uint64_t max_iter) {
    if (s) (s->csr ? s->csr->options() : s->impl->options()).max_iter = max_iter;
}

void fem_solver_set_tol(fem_solver* s, double tol) {
    if (s) (s->csr ? s->csr->options() : s->impl->options()).tol = tol;
}

void fem_solver_free(fem_solver* s) { delete s; }

uint64_t fem_solver_free_count(const fem_solver* s) {
    if (!s) return 0;
    return s->csr ? s->csr->n() : s->impl->model().free.size();
}

uint64_t fem_solver_unknown_count(const fem_solver* s) {
    if (!s) return 0;
    if (s->box) return s->box->unknowns();
    return s->csr ? s->csr->n() : s->impl->model().unknowns();
}

int fem_solver_free_list(const fem_solver* s, int64_t* out) {
    if (!s || !out) return 1;
    if (s->box) {
        std::memcpy(out, s->box->free.data(), s->box->free.size() * sizeof(int64_t));
        return 0;
    }
    if (s->csr) {  // the matrix's unknowns are all free, in its order
        for (uint64_t i = 0; i < s->csr->n(); ++i) out[i] = (int64_t)i;
        return 0;
    }
    const auto& f = s->impl->model().free;
    std::memcpy(out, f.data(), f.size() * sizeof(int64_t));
    return 0;
}

int fem_solver_point_grid(const fem_solver* s, uint32_t* out) {
    if (!s || !out || (s->csr && !s->box)) return 1;
    const auto& m = s->box ? *s->box : s->impl->model();
    for (size_t p = 0; p < m.points; ++p) {
        const uint32_t g = m.point_grid[p];
        out[3 * p] = g % m.gn[0];
        out[3 * p + 1] = (g / m.gn[0]) % m.gn[1];
        out[3 * p + 2] = g / (m.gn[0] * m.gn[1]);
    }
    return 0;
}

int fem_solver_rhs(fem_solver* s, double* out, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->box) {
            std::copy(s->box_rhs.begin(), s->box_rhs.end(), out);
            return;
        }
        if (s->csr) throw fem::Error("an assembled solver has no model: its right side is the caller's");
        s->impl->rhs(out);
    });
}

int fem_solver_apply(fem_solver* s, const double* x, double* y, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) s->csr->apply_host(x, y);
        else s->impl->apply_host(x, y);
    });
}

int fem_solver_precondition(fem_solver* s, const double* r, double* z, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->csr) s->csr->precondition_host(r, z);
        else s->impl->precondition_host(r, z);
    });
}

int fem_solver_matrix(fem_solver* s, uint64_t* nnz, uint64_t* ptr, uint32_t* idx, double* val, char* err,
                      size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!s) throw fem::Error("no solver");
        if (s->box) {  // the assembled matrix of the box
            const auto& a = s->box_matrix;
            if (ptr) std::copy(a.ptr.begin(), a.ptr.end(), ptr);
            if (idx) std::copy(a.idx.begin(), a.idx.end(), idx);
            if (val) std::copy(a.val.begin(), a.val.end(), val);
            if (nnz) *nnz = a.nnz();
            return;
        }
        if (s->csr) throw fem::Error("an assembled solver holds the caller's matrix already");
        const uint64_t n = s->impl->export_matrix(ptr, idx, val);
        if (nnz) *nnz = n;
    });
}

int fem_solver_level_info(const fem_solver* s, int32_t level, uint64_t* unknowns, double* lmax) {
    if (!s || !unknowns || !lmax) return 1;
    try {
        if (s->csr) s->csr->level_info(level, unknowns, lmax);
        else s->impl->level_info(level, unknowns, lmax);
        return 0;
    } catch (...) {
        return 1;
    }
}

const char* fem_solver_device_name(const fem_solver* s) {
    if (!s) return "";
    return s->csr ? s->csr->device_name() : s->impl->device_name();
}

void fem_solver_set_profile(fem_solver* s, int on) {
    if (s && s->impl) s->impl->set_profile(on != 0);
}

int fem_solver_time_kernels(fem_solver* s, int reps, fem_profile* out) {
    if (!s || !out || !s->impl) return 1;
    try {
        s->impl->time_kernels(reps, out);
        return 0;
    } catch (...) {
        return 1;
    }
}

int fem_solver_profile(fem_solver* s, fem_profile* out) {
    if (!s || !out || !s->impl) return 1;
    try {
        s->impl->profile(out);
        return 0;
    } catch (...) {
        return 1;
    }
}

struct fem_server {
    std::unique_ptr<fem::Server> impl;
};

int fem_server_new(const fem_server_options* opt, fem_server** out, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!opt || !out) throw fem::Error("fem_server_new needs options and an output handle");
        auto* srv = new fem_server;
        srv->impl = std::make_unique<fem::Server>(opt->hbm_budget, opt->max_active, opt->device, opt->setup_threads,
                                                  opt->unbatched == 0);
        *out = srv;
    });
}

int fem_server_submit(fem_server* srv, const fem_box* box, const fem_options* opt, const double* b, uint64_t* id,
                      char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!srv || !box || !id) throw fem::Error("fem_server_submit needs a server, a box and an id output");
        fem_options def;
        fem_options_default(&def);
        *id = srv->impl->submit(box_input(box), options_from(opt ? *opt : def), b);
    });
}

int fem_server_wait(fem_server* srv, uint64_t id, double timeout_s, double* x, uint64_t n, fem_stats* st,
                    fem_request_info* info, char* err, size_t errlen) {
    if (!srv) {
        put_err(err, errlen, "no server");
        return 1;
    }
    std::string e;
    const int rc = srv->impl->wait(id, timeout_s, x, n, st, info, e);
    if (rc == 1) put_err(err, errlen, e.c_str());
    return rc;
}

int fem_server_cancel(fem_server* srv, uint64_t id) { return srv && srv->impl->cancel(id) ? 0 : 1; }

int fem_server_memory(fem_server* srv, uint64_t* arena_bytes, uint64_t* peak_bytes, uint64_t* misses) {
    if (!srv) return 1;
    if (arena_bytes) *arena_bytes = srv->impl->arena_bytes();
    if (peak_bytes) *peak_bytes = srv->impl->arena_peak();
    if (misses) *misses = srv->impl->arena_misses();
    return 0;
}

int fem_server_batching(fem_server* srv, fem_server_batch_stats* out) {
    if (!srv || !out) return 1;
    srv->impl->batch_stats(out);
    return 0;
}

void fem_server_free(fem_server* srv) { delete srv; }

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// The model alone (host only).
struct fem_model {
    fem::BoxModel m;
};

namespace {
fem::BoxInput box_input(const fem_box* box) {
    if (!box) throw fem::Error("a box is needed");
    fem::BoxInput in;
    if (box->kind < 0 || box->kind > 4) throw fem::Error("unknown model kind");
    in.kind = (fem::Kind)box->kind;
    for (int d = 0; d < 3; ++d) {
        in.n[d] = box->n[d];
        in.len[d] = box->length[d];
    }
    if (box->n[0] < 1 || box->n[1] < 1 || box->n[2] < 1) throw fem::Error("each axis needs 1 cell or more");
    const size_t cells = (size_t)box->n[0] * box->n[1] * box->n[2];
    const size_t count = fem::hex_kind(in.kind) ? cells : 6 * cells;
    if (!box->coef0) throw fem::Error("the material values are missing");
    in.coef0.assign(box->coef0, box->coef0 + count);
    if (in.kind == fem::Kind::ElasticTet4 || in.kind == fem::Kind::ElasticTet10 || in.kind == fem::Kind::ElasticHex8) {
        if (!box->coef1) throw fem::Error("the Poisson ratios are missing");
        in.coef1.assign(box->coef1, box->coef1 + count);
    }
    for (int f = 0; f < 6; ++f) {
        in.fixed[f] = box->fixed[f] != 0;
        in.value[f] = box->value[f];
        for (int d = 0; d < 3; ++d) in.traction[f][d] = box->traction[3 * f + d];
    }
    in.source = box->source;
    for (int d = 0; d < 3; ++d) in.body[d] = box->body[d];
    return in;
}
}  // namespace

extern "C" {

int fem_test_hex_element(const double* size, double e, double nu, double* out, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!size || !out) throw fem::Error("the sizes and the output are needed");
        const auto k = fem::hex_element({size[0], size[1], size[2]}, fem::isotropic_stiffness(e, nu));
        std::copy(k.begin(), k.end(), out);
    });
}

int fem_hex_element(const double* size, const double* material, const double* alpha, double* k, double* load,
                    char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!size || !material || !k) throw fem::Error("the sizes, the material and the stiffness output are needed");
        std::array<double, 576> kk;
        std::array<double, 192> ll;
        const std::array<double, 3> a = alpha ? std::array<double, 3>{alpha[0], alpha[1], alpha[2]}
                                              : std::array<double, 3>{0.0, 0.0, 0.0};
        fem::hex_element_thermal({size[0], size[1], size[2]}, fem::orthotropic_stiffness(material), a, kk, ll);
        std::copy(kk.begin(), kk.end(), k);
        if (load) std::copy(ll.begin(), ll.end(), load);
    });
}

int fem_hex_stress(const double* size, const double* material, const double* alpha, const double* xi,
                   const double* u, const double* t, double* strain, double* stress, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!size || !material || !alpha || !xi || !u || !t || !strain || !stress)
            throw fem::Error("every argument is needed");
        fem::hex_strain_stress({size[0], size[1], size[2]}, fem::orthotropic_stiffness(material),
                               {alpha[0], alpha[1], alpha[2]}, xi, u, t, strain, stress);
    });
}

int fem_test_hex_box_system(const fem_box* box, uint64_t* n, uint64_t* nnz, uint64_t* ptr, uint32_t* idx, double* val,
                            double* rhs, double* aspect, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        const fem::BoxInput in = box_input(box);
        if (in.kind != fem::Kind::ElasticHex8) throw fem::Error("a hex8 elasticity box is needed");
        const fem::HexBoxSystem sys = fem::hex_box_system(fem::build_model(in));
        if (n) *n = sys.a.n;
        if (nnz) *nnz = sys.a.nnz();
        if (ptr) std::copy(sys.a.ptr.begin(), sys.a.ptr.end(), ptr);
        if (idx) std::copy(sys.a.idx.begin(), sys.a.idx.end(), idx);
        if (val) std::copy(sys.a.val.begin(), sys.a.val.end(), val);
        if (rhs) std::copy(sys.rhs.begin(), sys.rhs.end(), rhs);
        if (aspect) *aspect = sys.aspect;
    });
}

int fem_model_new_box(const fem_box* box, fem_model** out, char* err, size_t errlen) {
    return guarded(err, errlen, [&] {
        if (!out) throw fem::Error("an output handle is needed");
        *out = nullptr;
        auto* h = new fem_model;
        try {
            h->m = fem::build_model(box_input(box));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

uint64_t fem_model_point_count(const fem_model* h) { return h ? h->m.points : 0; }
uint64_t fem_model_unknown_count(const fem_model* h) { return h ? h->m.unknowns() : 0; }
uint64_t fem_model_free_count(const fem_model* h) { return h ? h->m.free.size() : 0; }

int fem_model_free_list(const fem_model* h, int64_t* out) {
    if (!h || !out) return 1;
    std::memcpy(out, h->m.free.data(), h->m.free.size() * sizeof(int64_t));
    return 0;
}

int fem_model_point_grid(const fem_model* h, uint32_t* out) {
    if (!h || !out) return 1;
    const auto& m = h->m;
    for (size_t p = 0; p < m.points; ++p) {
        const uint32_t g = m.point_grid[p];
        out[3 * p] = g % m.gn[0];
        out[3 * p + 1] = (g / m.gn[0]) % m.gn[1];
        out[3 * p + 2] = g / (m.gn[0] * m.gn[1]);
    }
    return 0;
}

int fem_model_load(const fem_model* h, double* out) {
    if (!h || !out) return 1;
    const auto& m = h->m;
    for (size_t p = 0; p < m.points; ++p)
        for (int c = 0; c < m.comps; ++c) out[m.comps * p + c] = m.load[(size_t)m.comps * m.point_grid[p] + c];
    return 0;
}

int fem_model_tet_cells(const fem_model* h, uint32_t* cell, uint8_t* type) {
    if (!h || !cell || !type || h->m.tet_cell.empty()) return 1;
    std::copy(h->m.tet_cell.begin(), h->m.tet_cell.end(), cell);
    std::copy(h->m.tet_type.begin(), h->m.tet_type.end(), type);
    return 0;
}

void fem_model_free(fem_model* h) { delete h; }

}  // extern "C"
