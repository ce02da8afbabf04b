// The host setup of the assembled (CSR) boundary: validation, the reference's grid coarsening,
// the Galerkin product and the level loop. No FMA contraction in this file: every product and
// sum is rounded where the reference's Rust rounds it (Rust does not fuse a * b + c).
#include "csr_host.hpp"

#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <limits>
#include <thread>
#include <utility>

#include "common.hpp"

namespace fem {
namespace csr {

namespace {
void check_rows(uint64_t rows, uint64_t cols, const std::vector<uint64_t>& ptr, const std::vector<uint32_t>& idx,
                size_t nval, const char* what) {
    const std::string w(what);
    if (ptr.size() != rows + 1 || ptr[0] != 0) throw Error(w + ": ptr needs rows + 1 offsets from 0");
    if (idx.size() != ptr[rows] || nval != ptr[rows]) throw Error(w + ": idx and val need ptr[rows] entries");
    for (uint64_t i = 0; i < rows; ++i) {
        if (ptr[i + 1] < ptr[i]) throw Error(w + ": ptr decreases");
        for (uint64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            if (idx[k] >= cols) throw Error(w + ": a column index is out of range");
            if (k > ptr[i] && idx[k] <= idx[k - 1]) throw Error(w + ": the columns of a row are not increasing");
        }
    }
}
}  // namespace

void validate(const Csr& a) {
    if (a.n == 0) throw Error("the matrix has no unknowns");
    if (a.n > std::numeric_limits<uint32_t>::max()) throw Error("the matrix has more than 2^32 - 1 unknowns");
    check_rows(a.n, a.n, a.ptr, a.idx, a.val.size(), "the matrix");
    for (uint64_t i = 0; i < a.n; ++i) {
        const auto b = a.idx.begin() + a.ptr[i], e = a.idx.begin() + a.ptr[i + 1];
        if (!std::binary_search(b, e, (uint32_t)i)) throw Error("the matrix has a row without its diagonal entry");
    }
}

void validate(const Rect& p) { check_rows(p.rows, p.cols, p.ptr, p.idx, p.val.size(), "the prolongation"); }

void validate(const GridLevel& g, uint64_t n) {
    if (g.points.size() != n || g.comp.size() != n) throw Error("the grid level needs a point and a component for each unknown");
    for (int d = 0; d < 3; ++d) {
        if (g.planes[d].empty()) throw Error("each axis of the grid needs a plane");
        for (size_t i = 1; i < g.planes[d].size(); ++i)
            if (!(g.planes[d][i] > g.planes[d][i - 1])) throw Error("the planes of an axis must increase");
    }
    for (uint64_t u = 0; u < n; ++u) {
        for (int d = 0; d < 3; ++d)
            if (g.points[u][d] >= g.planes[d].size()) throw Error("a grid point is outside the planes");
        if (g.comp[u] > 2) throw Error("a component is above 2");
    }
}

bool coarsen(const GridLevel& fine, const Csr& a, Rect& P, GridLevel& coarse) {
    // Per axis, per fine plane: two (coarse plane, weight) pairs; a kept plane maps to itself with
    // weights 1 and 0.
    struct Wt {
        uint32_t c;
        double w;
    };
    std::array<std::vector<std::array<Wt, 2>>, 3> map;
    bool any = false;
    for (int d = 0; d < 3; ++d) {
        const std::vector<double>& pl = fine.planes[d];
        const size_t n = pl.size();
        std::vector<size_t> keep;
        if (n <= 2) {
            for (size_t i = 0; i < n; ++i) keep.push_back(i);
        } else {
            any = true;
            for (size_t i = 0; i < n; i += 2) keep.push_back(i);
            if (keep.back() != n - 1) keep.push_back(n - 1);
        }
        coarse.planes[d].clear();
        for (size_t i : keep) coarse.planes[d].push_back(pl[i]);
        map[d].resize(n);
        size_t c = 0;
        for (size_t i = 0; i < n; ++i) {
            while (keep[c] < i) ++c;
            if (keep[c] == i) {
                map[d][i] = {Wt{(uint32_t)c, 1.0}, Wt{(uint32_t)c, 0.0}};
            } else {
                const size_t lo = keep[c - 1], hi = keep[c];
                const double w = (pl[i] - pl[lo]) / (pl[hi] - pl[lo]);
                map[d][i] = {Wt{(uint32_t)(c - 1), 1.0 - w}, Wt{(uint32_t)c, w}};
            }
        }
    }
    if (!any) return false;
    const uint64_t dims[3] = {coarse.planes[0].size(), coarse.planes[1].size(), coarse.planes[2].size()};
    uint32_t cmax = 0;
    for (uint8_t c : fine.comp) cmax = std::max<uint32_t>(cmax, c);
    const uint64_t comps = fine.comp.empty() ? 1 : cmax + 1;
    std::vector<uint32_t> number(comps * dims[0] * dims[1] * dims[2], UINT32_MAX);
    coarse.points.clear();
    coarse.comp.clear();
    P.rows = fine.points.size();
    P.ptr.assign(1, 0);
    P.idx.clear();
    P.val.clear();
    std::vector<std::pair<uint32_t, double>> row;
    for (uint64_t r = 0; r < fine.points.size(); ++r) {
        const auto& f = fine.points[r];
        const uint8_t c = fine.comp[r];
        std::array<std::array<Wt, 2>, 3> wts;
        for (int d = 0; d < 3; ++d) {
            std::array<Wt, 2> w = map[d][f[d]];
            if (w[0].c != w[1].c && comps == 1) {
                // operator weights: this row's couplings to the plane below and the plane above
                double lo = 0.0, hi = 0.0;
                for (uint64_t k = a.ptr[r]; k < a.ptr[r + 1]; ++k) {
                    const auto& g = fine.points[a.idx[k]];
                    if (g[d] + 1 == f[d]) lo -= a.val[k];
                    else if (g[d] == f[d] + 1) hi -= a.val[k];
                }
                if (lo > 0.0 && hi > 0.0) {
                    w[0].w = lo / (lo + hi);
                    w[1].w = hi / (lo + hi);
                }
            }
            wts[d] = w;
        }
        row.clear();
        for (const Wt& x : wts[0])
            for (const Wt& y : wts[1])
                for (const Wt& z : wts[2]) {
                    const double w = x.w * y.w * z.w;
                    if (w == 0.0) continue;
                    const uint64_t slot = c + comps * (x.c + dims[0] * (y.c + dims[1] * (uint64_t)z.c));
                    if (number[slot] == UINT32_MAX) {
                        number[slot] = (uint32_t)coarse.points.size();
                        coarse.points.push_back({x.c, y.c, z.c});
                        coarse.comp.push_back(c);
                    }
                    row.emplace_back(number[slot], w);
                }
        std::sort(row.begin(), row.end(), [](const auto& p, const auto& q) { return p.first < q.first; });
        for (const auto& e : row) {
            P.idx.push_back(e.first);
            P.val.push_back(e.second);
        }
        P.ptr.push_back(P.idx.size());
    }
    P.cols = coarse.points.size();
    return true;
}

Rect transpose(const Rect& p) {
    Rect t;
    t.rows = p.cols;
    t.cols = p.rows;
    t.ptr.assign(p.cols + 1, 0);
    for (uint32_t j : p.idx) ++t.ptr[j + 1];
    for (uint64_t j = 0; j < p.cols; ++j) t.ptr[j + 1] += t.ptr[j];
    t.idx.resize(p.idx.size());
    t.val.resize(p.idx.size());
    std::vector<uint64_t> next(t.ptr.begin(), t.ptr.end() - 1);
    for (uint64_t i = 0; i < p.rows; ++i)
        for (uint64_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k) {
            const uint64_t at = next[p.idx[k]]++;
            t.idx[at] = (uint32_t)i;
            t.val[at] = p.val[k];
        }
    return t;
}

Csr galerkin(const Rect& p, const Csr& a, int threads) {
    const uint64_t nc = p.cols;
    const Rect t = transpose(p);
    // Each thread builds whole rows into its own lists; the rows are then concatenated in order.
    const int nt = std::max(1, std::min<int>(threads, (int)std::min<uint64_t>(nc, 64)));
    std::vector<std::vector<uint64_t>> rlen(nt);
    std::vector<std::vector<uint32_t>> ridx(nt);
    std::vector<std::vector<double>> rval(nt);
    auto work = [&](int id) {
        const uint64_t r0 = nc * id / nt, r1 = nc * (id + 1) / nt;
        std::vector<double> acc(nc, 0.0);
        std::vector<uint64_t> mark(nc, UINT64_MAX);
        std::vector<uint32_t> row;
        for (uint64_t c = r0; c < r1; ++c) {
            row.clear();
            for (uint64_t q = t.ptr[c]; q < t.ptr[c + 1]; ++q) {
                const uint64_t i = t.idx[q];
                const double pic = t.val[q];
                for (uint64_t ka = a.ptr[i]; ka < a.ptr[i + 1]; ++ka) {
                    const uint64_t l = a.idx[ka];
                    const double aval = a.val[ka] * pic;
                    for (uint64_t kp = p.ptr[l]; kp < p.ptr[l + 1]; ++kp) {
                        const uint32_t d = p.idx[kp];
                        if (mark[d] != c) {
                            mark[d] = c;
                            acc[d] = 0.0;
                            row.push_back(d);
                        }
                        acc[d] += aval * p.val[kp];
                    }
                }
            }
            std::sort(row.begin(), row.end());
            rlen[id].push_back(row.size());
            for (uint32_t d : row) {
                ridx[id].push_back(d);
                rval[id].push_back(acc[d]);
            }
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int id = 0; id < nt; ++id) pool.emplace_back(work, id);
        for (auto& th : pool) th.join();
    }
    Csr out;
    out.n = nc;
    out.ptr.reserve(nc + 1);
    out.ptr.push_back(0);
    for (int id = 0; id < nt; ++id) {
        for (uint64_t len : rlen[id]) out.ptr.push_back(out.ptr.back() + len);
        out.idx.insert(out.idx.end(), ridx[id].begin(), ridx[id].end());
        out.val.insert(out.val.end(), rval[id].begin(), rval[id].end());
    }
    return out;
}

std::vector<double> diagonal(const Csr& a) {
    std::vector<double> d(a.n, 0.0);
    for (uint64_t i = 0; i < a.n; ++i) {
        const auto b = a.idx.begin() + a.ptr[i], e = a.idx.begin() + a.ptr[i + 1];
        const auto it = std::lower_bound(b, e, (uint32_t)i);
        if (it != e && *it == i) d[i] = a.val[a.ptr[i] + (it - b)];
    }
    return d;
}

GridHierarchy grid_hierarchy(const Csr& a0, const GridLevel& level0, uint64_t coarsest, int max_levels, int threads) {
    GridHierarchy h;
    h.a.push_back(a0);
    h.level.push_back(level0);
    for (;;) {
        const Csr& a = h.a.back();
        if (a.n <= coarsest || h.p.size() >= 30) break;
        if (max_levels > 0 && (int)h.a.size() >= max_levels) break;
        Rect p;
        GridLevel c;
        if (!coarsen(h.level.back(), a, p, c)) break;
        if (c.points.size() * 10 > a.n * 9) break;
        Csr ac = galerkin(p, a, threads);
        h.p.push_back(std::move(p));
        h.level.push_back(std::move(c));
        h.a.push_back(std::move(ac));
    }
    return h;
}

}  // namespace csr
}  // namespace fem
