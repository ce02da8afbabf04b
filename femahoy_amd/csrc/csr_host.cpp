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

bool coarsen(const GridLevel& fine, const Csr& a, Rect& P, GridLevel& coarse, std::array<std::vector<size_t>, 3>* kept) {
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
        if (kept) (*kept)[d] = keep;
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

// hex.rs:164-178: the 8 trilinear corner functions at the local point xi, corners in BITS order.
static std::array<double, 8> hex_shape(const double xi[3]) {
    static const int BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    std::array<double, 8> out{};
    for (int n = 0; n < 8; ++n) {
        double p = 1.0;
        for (int d = 0; d < 3; ++d) p *= BITS[n][d] == 1 ? 0.5 * (1.0 + xi[d]) : 0.5 * (1.0 - xi[d]);
        out[n] = p;
    }
    return out;
}

// multigrid.rs:609-643: the modes of a coarse cell to keep (3 d + c), by a Cholesky factor of the
// Gram matrix of the 11 functions at the free fine unknowns of each component, corners first: a
// mode whose pivot is at most 1e-6 of its diagonal is left out.
static std::array<bool, 9> coarse_modes(const double gram[3][11][11]) {
    std::array<bool, 9> keep{};
    for (int c = 0; c < 3; ++c) {
        double l[11][11] = {};
        bool used[11] = {};
        for (int j = 0; j < 11; ++j) {
            double d = gram[c][j][j];
            for (int k = 0; k < j; ++k)
                if (used[k]) d -= l[j][k] * l[j][k];
            if (!(d > 1e-6 * gram[c][j][j])) continue;
            used[j] = true;
            d = std::sqrt(d);
            l[j][j] = d;
            for (int i = j + 1; i < 11; ++i) {
                double v = gram[c][i][j];
                for (int k = 0; k < j; ++k)
                    if (used[k]) v -= l[i][k] * l[j][k];
                l[i][j] = v / d;
            }
            if (j >= 8) keep[3 * (j - 8) + c] = true;
        }
    }
    return keep;
}

// multigrid.rs:648-671: the 24 x 24 corner matrix by symmetric Gauss elimination of the kept modes
// of the 33 x 33 matrix k (changed); a mode with no energy is left out too.
static std::array<double, 576> condense_modes(std::vector<double>& k, const std::array<bool, 9>& keep) {
    for (int m = 24; m < 33; ++m) {
        if (!keep[m - 24]) continue;
        const double pivot = k[m * 33 + m];
        double scale = 0.0;
        for (int i = 0; i < 33; ++i) scale = std::max(scale, std::abs(k[i * 33 + i]));
        if (!(pivot > 1e-12 * scale)) continue;
        for (int i = 0; i < 33; ++i) {
            const double f = k[i * 33 + m] / pivot;
            if (i == m || f == 0.0) continue;
            for (int j = 0; j < 33; ++j) k[i * 33 + j] -= f * k[m * 33 + j];
        }
    }
    std::array<double, 576> out{};
    for (int r = 0; r < 24; ++r)
        for (int c = 0; c < 24; ++c) out[r * 24 + c] = k[r * 33 + c];
    return out;
}

Csr by_cells(const GridLevel& fine, const Cells& fc, const GridLevel& coarse, const std::array<std::vector<size_t>, 3>& keep,
             Cells& next) {
    static const uint32_t BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const uint64_t fd[3] = {fine.planes[0].size(), fine.planes[1].size(), fine.planes[2].size()};
    const uint64_t cd[3] = {coarse.planes[0].size(), coarse.planes[1].size(), coarse.planes[2].size()};
    auto index = [](const std::array<uint32_t, 3>& q, uint64_t c, const uint64_t* d) {
        return c + 3 * (q[0] + d[0] * (q[1] + d[1] * (uint64_t)q[2]));
    };
    std::vector<uint8_t> free(3 * fd[0] * fd[1] * fd[2], 0);
    for (size_t k = 0; k < fine.points.size(); ++k) free[index(fine.points[k], fine.comp[k], fd)] = 1;
    std::vector<uint32_t> number(3 * cd[0] * cd[1] * cd[2], UINT32_MAX);
    for (size_t k = 0; k < coarse.points.size(); ++k) number[index(coarse.points[k], coarse.comp[k], cd)] = (uint32_t)k;
    // the coarse cell (interval of kept planes) of each fine plane along each axis
    std::array<std::vector<uint32_t>, 3> interval;
    for (int d = 0; d < 3; ++d) {
        size_t c = 0;
        for (uint64_t i = 0; i < fd[d]; ++i) {
            while (c + 2 < keep[d].size() && keep[d][c + 1] <= i) ++c;
            interval[d].push_back((uint32_t)c);
        }
    }
    // fine cells grouped by coarse cell (z, y, x); within a coarse cell in increasing fine index
    std::vector<std::pair<uint32_t, std::array<uint32_t, 3>>> order(fc.origin.size());
    for (size_t f = 0; f < fc.origin.size(); ++f) {
        const auto& o = fc.origin[f];
        order[f] = {(uint32_t)f, {interval[0][o[0]], interval[1][o[1]], interval[2][o[2]]}};
    }
    std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) {
        if (x.second[2] != y.second[2]) return x.second[2] < y.second[2];
        if (x.second[1] != y.second[1]) return x.second[1] < y.second[1];
        return x.second[0] < y.second[0];
    });
    next.origin.clear();
    next.k.clear();
    size_t start = 0;
    std::vector<double> k(33 * 33);
    while (start < order.size()) {
        const auto cc = order[start].second;
        size_t end = start;
        while (end < order.size() && order[end].second == cc) ++end;
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = coarse.planes[d][cc[d]];
            hi[d] = coarse.planes[d][cc[d] + 1];
        }
        std::fill(k.begin(), k.end(), 0.0);
        double gram[3][11][11] = {};
        for (size_t t = start; t < end; ++t) {
            const uint32_t f = order[t].first;
            const auto& kf = fc.k[f];
            const auto& o = fc.origin[f];
            // the 11 coarse functions at the 8 fine corners, 0 at a fixed fine unknown
            double phi[8][3][11] = {};
            for (int n = 0; n < 8; ++n) {
                const std::array<uint32_t, 3> q{o[0] + BITS[n][0], o[1] + BITS[n][1], o[2] + BITS[n][2]};
                double xi[3];
                for (int d = 0; d < 3; ++d) xi[d] = 2.0 * (fine.planes[d][q[d]] - lo[d]) / (hi[d] - lo[d]) - 1.0;
                const auto corner = hex_shape(xi);
                for (int c = 0; c < 3; ++c) {
                    if (!free[index(q, c, fd)]) continue;
                    double* fv = phi[n][c];
                    for (int a = 0; a < 8; ++a) fv[a] = corner[a];
                    for (int d = 0; d < 3; ++d) fv[8 + d] = 1.0 - xi[d] * xi[d];
                    for (int a = 0; a < 11; ++a)
                        for (int b = 0; b < 11; ++b) gram[c][a][b] += fv[a] * fv[b];
                }
            }
            // kt = Kf T, then K += T^T kt
            double kt[24][33] = {};
            for (int r = 0; r < 24; ++r)
                for (int m = 0; m < 8; ++m)
                    for (int c = 0; c < 3; ++c) {
                        const double v = kf[r * 24 + 3 * m + c];
                        if (v != 0.0)
                            for (int a = 0; a < 11; ++a) kt[r][3 * a + c] += v * phi[m][c][a];
                    }
            for (int n = 0; n < 8; ++n)
                for (int c = 0; c < 3; ++c) {
                    const double* row = kt[3 * n + c];
                    for (int a = 0; a < 11; ++a) {
                        const double w = phi[n][c][a];
                        if (w == 0.0) continue;
                        double* out = &k[(3 * a + c) * 33];
                        for (int j = 0; j < 33; ++j) out[j] += w * row[j];
                    }
                }
        }
        next.k.push_back(condense_modes(k, coarse_modes(gram)));
        next.origin.push_back(cc);
        start = end;
    }
    // assembly on the coarse unknowns (Csr::pattern then add_element, sparse.rs:22-86)
    const uint64_t nc = coarse.points.size();
    std::vector<std::vector<std::pair<int, uint32_t>>> ids(next.origin.size());
    for (size_t e = 0; e < next.origin.size(); ++e) {
        const auto& o = next.origin[e];
        for (int n = 0; n < 8; ++n) {
            const std::array<uint32_t, 3> q{o[0] + BITS[n][0], o[1] + BITS[n][1], o[2] + BITS[n][2]};
            for (int c = 0; c < 3; ++c) {
                const uint32_t g = number[index(q, c, cd)];
                if (g != UINT32_MAX) ids[e].push_back({3 * n + c, g});
            }
        }
    }
    std::vector<std::vector<uint32_t>> rows(nc);
    for (uint64_t i = 0; i < nc; ++i) rows[i].push_back((uint32_t)i);
    for (const auto& list : ids)
        for (const auto& r : list)
            for (const auto& c : list) rows[r.second].push_back(c.second);
    Csr A;
    A.n = nc;
    A.ptr.assign(nc + 1, 0);
    for (uint64_t i = 0; i < nc; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
        A.ptr[i + 1] = A.ptr[i] + rows[i].size();
    }
    A.idx.reserve(A.ptr[nc]);
    for (const auto& r : rows) A.idx.insert(A.idx.end(), r.begin(), r.end());
    A.val.assign(A.ptr[nc], 0.0);
    for (size_t e = 0; e < ids.size(); ++e)
        for (const auto& r : ids[e])
            for (const auto& c : ids[e]) {
                const uint32_t* b = A.idx.data() + A.ptr[r.second];
                const uint32_t* f = std::lower_bound(b, (const uint32_t*)A.idx.data() + A.ptr[r.second + 1], c.second);
                A.val[(uint64_t)(f - (const uint32_t*)A.idx.data())] += next.k[e][r.first * 24 + c.first];
            }
    return A;
}

void validate(const Cells& cells, const GridLevel& g) {
    if (cells.k.size() != cells.origin.size()) throw Error("one matrix for each cell");
    for (const auto& o : cells.origin)
        for (int d = 0; d < 3; ++d)
            if ((size_t)o[d] + 1 >= g.planes[d].size()) throw Error("a cell's corner is outside the planes");
    for (uint8_t c : g.comp)
        if (c > 2) throw Error("cells need displacement unknowns (components 0 to 2)");
}

GridHierarchy grid_hierarchy(const Csr& a0, const GridLevel& level0, uint64_t coarsest, int max_levels, int threads,
                             const Cells* cells) {
    GridHierarchy h;
    h.a.push_back(a0);
    h.level.push_back(level0);
    Cells stored;
    bool have_stored = false;
    for (;;) {
        const Csr& a = h.a.back();
        if (a.n <= coarsest || h.p.size() >= 30) break;
        if (max_levels > 0 && (int)h.a.size() >= max_levels) break;
        Rect p;
        GridLevel c;
        std::array<std::vector<size_t>, 3> kept;
        if (!coarsen(h.level.back(), a, p, c, &kept)) break;
        if (c.points.size() * 10 > a.n * 9) break;
        Csr ac;
        if (cells) {
            // multigrid.rs:109-118: the coarse matrices from the cells of the level
            Cells next;
            ac = by_cells(h.level.back(), have_stored ? stored : *cells, c, kept, next);
            stored = std::move(next);
            have_stored = true;
        } else {
            ac = galerkin(p, a, threads);
        }
        h.p.push_back(std::move(p));
        h.level.push_back(std::move(c));
        h.a.push_back(std::move(ac));
    }
    return h;
}

}  // namespace csr
}  // namespace fem
