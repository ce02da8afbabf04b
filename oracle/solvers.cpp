// TEST INFRASTRUCTURE ONLY (the oracle). Restates crates/fem/src/multigrid.rs (pcg, Lines,
// coarsen, Multigrid for a scalar), pmultigrid.rs and solve.rs.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <tuple>

#include "oracle.hpp"

namespace orc {

static const usize COARSEST = 3000;  // multigrid.rs:50
static const usize SWEEPS = 2;       // pmultigrid.rs:33

static double dot(const double* u, const double* v, usize n) {
    double s = 0.0;
    for (usize i = 0; i < n; ++i) s += u[i] * v[i];
    return s;
}

// Rust's `{:.2e}`: mantissa with 2 decimals, exponent with no '+' and no leading zeros.
static std::string rust_e2(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.2e", v);
    std::string t(buf);
    const usize epos = t.find('e');
    if (epos == std::string::npos) return t;
    return t.substr(0, epos) + "e" + std::to_string(std::atoi(t.c_str() + epos + 1));
}

// multigrid.rs:692-753
Convergence pcg(const Csr& a, const double* b, double* x, const Precondition& precondition, double tol,
                usize max_iter, const Cancelled& cancelled) {
    const usize n = a.n;
    const double norm_b = std::sqrt(dot(b, b, n));
    if (norm_b == 0.0) {
        std::fill(x, x + n, 0.0);
        return {0, 0.0};
    }
    std::vector<double> r(n, 0.0);
    a.mul(x, r.data());
    for (usize i = 0; i < n; ++i) r[i] = b[i] - r[i];
    std::vector<double> z(n, 0.0);
    precondition(r.data(), z.data());
    std::vector<double> p = z;
    double rz = dot(r.data(), z.data(), n);
    std::vector<double> q(n, 0.0);
    for (usize it = 0; it <= max_iter; ++it) {
        const double residual = std::sqrt(dot(r.data(), r.data(), n)) / norm_b;
        if (residual <= tol) return {it, residual};
        if (it == max_iter)
            throw Error("conjugate gradients did not converge: the residual is " + rust_e2(residual) + " after " +
                        std::to_string(max_iter) + " iterations");
        if (cancelled && cancelled()) throw Error("cancelled");
        a.mul(p.data(), q.data());
        const double alpha = rz / dot(p.data(), q.data(), n);
        for (usize i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        precondition(r.data(), z.data());
        const double rz_new = dot(r.data(), z.data(), n);
        const double beta = rz_new / rz;
        rz = rz_new;
        for (usize i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    throw Error("the loop returns");
}

// ------------------------------------------------------------------ Lines, multigrid.rs:201-276
Lines Lines::make(const Level& level, usize axis, const Csr& a) {
    usize u, v;
    if (axis == 0) {
        u = 1;
        v = 2;
    } else if (axis == 1) {
        u = 0;
        v = 2;
    } else {
        u = 0;
        v = 1;
    }
    const usize n = level.points.size();
    Lines out;
    out.order.resize(n);
    for (usize k = 0; k < n; ++k) out.order[k] = (u32)k;
    auto key = [&](u32 k) {
        const auto& p = level.points[k];
        return std::make_tuple(p[v], p[u], p[axis], level.comp[k]);
    };
    // Keys are distinct (one unknown per point and component), so the unstable sort is unique.
    std::sort(out.order.begin(), out.order.end(), [&](u32 x, u32 y) { return key(x) < key(y); });
    bool have_last = false;
    std::array<u32, 3> last{};
    for (usize r = 0; r < n; ++r) {
        const auto& p = level.points[out.order[r]];
        const bool next = have_last && last[u] == p[u] && last[v] == p[v] &&
                          (last[axis] == p[axis] || last[axis] + 1 == p[axis]);
        if (!next) out.start.push_back(r);
        last = p;
        have_last = true;
    }
    out.start.push_back(n);
    out.line.assign(n, 0);
    out.place.assign(n, 0);
    for (usize l = 0; l + 1 < out.start.size(); ++l)
        for (usize r = out.start[l]; r < out.start[l + 1]; ++r) {
            out.line[out.order[r]] = (u32)l;
            out.place[out.order[r]] = (u32)r;
        }
    usize comps = 1;
    for (auto c : level.comp) comps = std::max(comps, (usize)c + 1);
    const usize band = 2 * comps - 1;
    const usize w = band + 1;
    out.band = band;
    out.factor.assign(n * w, 0.0);
    auto& factor = out.factor;
    for (usize l = 0; l + 1 < out.start.size(); ++l) {
        for (usize r = out.start[l]; r < out.start[l + 1]; ++r) {
            const usize i = out.order[r];
            const usize first = std::max(r >= band ? r - band : 0, out.start[l]);
            for (usize c = first; c <= r; ++c) {
                const usize e = a.find(i, out.order[c]);
                double s = e == NONE ? 0.0 : a.val[e];
                for (usize k = std::max(first, c >= band ? c - band : 0); k < c; ++k)
                    s -= factor[r * w + band + k - r] * factor[c * w + band + k - c];
                if (c == r) {
                    if (!(s > 0.0)) throw Error("the matrix is not positive definite: a part can move with no strain");
                    factor[r * w + band] = std::sqrt(s);
                } else {
                    factor[r * w + band + c - r] = s / factor[c * w + band];
                }
            }
        }
    }
    return out;
}

// multigrid.rs:280-324
void Lines::sweep(const Csr& a, const double* b, double* x, bool forward) const {
    const usize w = band + 1;
    std::vector<double> y;
    auto solve = [&](usize l) {
        const usize first = start[l], end = start[l + 1];
        y.clear();
        for (usize r = first; r < end; ++r) {
            const usize i = order[r];
            double s = b[i];
            for (usize k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
                const usize j = a.idx[k];
                const usize pj = place[j];
                const usize dist = pj > r ? pj - r : r - pj;
                const bool near = line[j] == l && dist <= band;
                if (!near) s -= a.val[k] * x[j];
            }
            y.push_back(s);
        }
        for (usize r = first; r < end; ++r) {
            double s = y[r - first];
            for (usize k = std::max(r >= band ? r - band : 0, first); k < r; ++k)
                s -= factor[r * w + band + k - r] * y[k - first];
            y[r - first] = s / factor[r * w + band];
        }
        for (usize r = end; r-- > first;) {
            double s = y[r - first];
            for (usize k = r + 1; k < std::min(r + w, end); ++k) s -= factor[k * w + band + r - k] * y[k - first];
            y[r - first] = s / factor[r * w + band];
        }
        for (usize r = first; r < end; ++r) x[order[r]] = y[r - first];
    };
    const usize lines = start.size() - 1;
    if (forward)
        for (usize l = 0; l < lines; ++l) solve(l);
    else
        for (usize l = lines; l-- > 0;) solve(l);
}

// ------------------------------------------------------------------ coarsen, multigrid.rs:329-452
std::optional<Coarsened> coarsen(const Level& fine, const Csr& a) {
    Coarsened out;
    std::array<std::vector<std::array<std::pair<u32, double>, 2>>, 3> map;
    bool any = false;
    for (usize d = 0; d < 3; ++d) {
        const auto& p = fine.planes[d];
        const usize n = p.size();
        std::vector<usize> keep;
        if (n <= 2) {
            for (usize i = 0; i < n; ++i) keep.push_back(i);
        } else {
            any = true;
            for (usize i = 0; i < n; i += 2) keep.push_back(i);
            if (keep.back() != n - 1) keep.push_back(n - 1);
        }
        for (usize i : keep) out.coarse.planes[d].push_back(p[i]);
        out.kept[d] = keep;
        usize c = 0;
        map[d].resize(n);
        for (usize i = 0; i < n; ++i) {
            while (keep[c] < i) c += 1;
            if (keep[c] == i) {
                map[d][i] = {std::make_pair((u32)c, 1.0), std::make_pair((u32)c, 0.0)};
            } else {
                const usize lo = keep[c - 1], hi = keep[c];
                const double w = (p[i] - p[lo]) / (p[hi] - p[lo]);
                map[d][i] = {std::make_pair((u32)(c - 1), 1.0 - w), std::make_pair((u32)c, w)};
            }
        }
    }
    if (!any) return std::nullopt;
    const usize dims[3] = {out.coarse.planes[0].size(), out.coarse.planes[1].size(), out.coarse.planes[2].size()};
    usize comps = 1;
    for (auto c : fine.comp) comps = std::max(comps, (usize)c + 1);
    auto flat = [&](const std::array<u32, 3>& q, std::uint8_t c) {
        return (usize)c + comps * ((usize)q[0] + dims[0] * ((usize)q[1] + dims[1] * (usize)q[2]));
    };
    std::vector<u32> number(comps * dims[0] * dims[1] * dims[2], UINT32_MAX);
    Rect& P = out.p;
    P.ptr.push_back(0);
    for (usize row = 0; row < fine.points.size(); ++row) {
        const usize start = P.idx.size();
        const auto& f = fine.points[row];
        const std::uint8_t c = fine.comp[row];
        std::array<std::array<std::pair<u32, double>, 2>, 3> weights;
        for (usize d = 0; d < 3; ++d) {
            auto w = map[d][f[d]];
            if (w[0].first != w[1].first && comps == 1) {
                double lo = 0.0, hi = 0.0;
                for (usize k = a.ptr[row]; k < a.ptr[row + 1]; ++k) {
                    const auto& g = fine.points[a.idx[k]];
                    if (g[d] + 1 == f[d]) lo -= a.val[k];
                    else if (g[d] == f[d] + 1) hi -= a.val[k];
                }
                if (lo > 0.0 && hi > 0.0) {
                    w[0].second = lo / (lo + hi);
                    w[1].second = hi / (lo + hi);
                }
            }
            weights[d] = w;
        }
        for (const auto& [cx, wx] : weights[0])
            for (const auto& [cy, wy] : weights[1])
                for (const auto& [cz, wz] : weights[2]) {
                    const double w = wx * wy * wz;
                    if (w == 0.0) continue;
                    const std::array<u32, 3> q = {cx, cy, cz};
                    u32& slot = number[flat(q, c)];
                    if (slot == UINT32_MAX) {
                        slot = (u32)out.coarse.points.size();
                        out.coarse.points.push_back(q);
                        out.coarse.comp.push_back(c);
                    }
                    P.idx.push_back(slot);
                    P.val.push_back(w);
                }
        // The row's columns are distinct (distinct coarse points), so the unstable sort is unique.
        std::vector<std::pair<u32, double>> pairs;
        for (usize k = start; k < P.idx.size(); ++k) pairs.emplace_back(P.idx[k], P.val[k]);
        std::sort(pairs.begin(), pairs.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (usize k = 0; k < pairs.size(); ++k) {
            P.idx[start + k] = pairs[k].first;
            P.val[start + k] = pairs[k].second;
        }
        P.ptr.push_back(P.idx.size());
    }
    P.rows = fine.points.size();
    P.cols = out.coarse.points.size();
    return out;
}

// ------------------------------------------------------------------ Multigrid, multigrid.rs:84-179
static std::vector<Vec3> level_xyz(const Level& level) {
    std::vector<Vec3> xyz(level.points.size());
    for (usize k = 0; k < level.points.size(); ++k)
        for (usize d = 0; d < 3; ++d) xyz[k][d] = level.planes[d][level.points[k][d]];
    return xyz;
}

// ------------------------------------------------------------------ by_cells, multigrid.rs:458-671
ByCells by_cells(const Level& fine, const CellMatrices& fc, const Level& coarse,
                 const std::array<std::vector<usize>, 3>& keep) {
    static const u32 BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const usize fd[3] = {fine.planes[0].size(), fine.planes[1].size(), fine.planes[2].size()};
    const usize cd[3] = {coarse.planes[0].size(), coarse.planes[1].size(), coarse.planes[2].size()};
    auto index = [](const std::array<u32, 3>& q, usize c, const usize* d) {
        return c + 3 * (q[0] + d[0] * (q[1] + d[1] * (usize)q[2]));
    };
    std::vector<bool> free(3 * fd[0] * fd[1] * fd[2], false);
    for (usize k = 0; k < fine.points.size(); ++k) free[index(fine.points[k], fine.comp[k], fd)] = true;
    std::vector<u32> number(3 * cd[0] * cd[1] * cd[2], UINT32_MAX);
    for (usize k = 0; k < coarse.points.size(); ++k) number[index(coarse.points[k], coarse.comp[k], cd)] = (u32)k;
    // the coarse cell of each fine cell along each axis
    std::array<std::vector<u32>, 3> interval;
    for (usize d = 0; d < 3; ++d) {
        usize c = 0;
        for (usize i = 0; i < fd[d]; ++i) {
            while (c + 2 < keep[d].size() && keep[d][c + 1] <= i) ++c;
            interval[d].push_back((u32)c);
        }
    }
    std::vector<std::pair<u32, std::array<u32, 3>>> order;
    for (usize f = 0; f < fc.origin.size(); ++f) {
        const auto& o = fc.origin[f];
        order.push_back({(u32)f, {interval[0][o[0]], interval[1][o[1]], interval[2][o[2]]}});
    }
    std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) {
        return std::make_tuple(x.second[2], x.second[1], x.second[0]) < std::make_tuple(y.second[2], y.second[1], y.second[0]);
    });
    ByCells out;
    usize start = 0;
    while (start < order.size()) {
        const auto cc = order[start].second;
        usize end = start;
        while (end < order.size() && order[end].second == cc) ++end;
        double lo[3], hi[3];
        for (usize d = 0; d < 3; ++d) {
            lo[d] = coarse.planes[d][cc[d]];
            hi[d] = coarse.planes[d][cc[d] + 1];
        }
        std::vector<double> k(33 * 33, 0.0);
        std::array<std::array<std::array<double, 11>, 11>, 3> gram{};
        for (usize t = start; t < end; ++t) {
            const u32 f = order[t].first;
            const auto& kf = fc.k[f];
            const auto& o = fc.origin[f];
            double phi[8][3][11] = {};
            for (usize n = 0; n < 8; ++n) {
                const std::array<u32, 3> q{o[0] + BITS[n][0], o[1] + BITS[n][1], o[2] + BITS[n][2]};
                Vec3 xi;
                for (usize d = 0; d < 3; ++d) xi[d] = 2.0 * (fine.planes[d][q[d]] - lo[d]) / (hi[d] - lo[d]) - 1.0;
                const auto corner = hex_shape(xi);
                for (usize c = 0; c < 3; ++c) {
                    if (!free[index(q, c, fd)]) continue;
                    double* fv = phi[n][c];
                    for (usize a = 0; a < 8; ++a) fv[a] = corner[a];
                    for (usize d = 0; d < 3; ++d) fv[8 + d] = 1.0 - xi[d] * xi[d];
                    for (usize a = 0; a < 11; ++a)
                        for (usize b = 0; b < 11; ++b) gram[c][a][b] += fv[a] * fv[b];
                }
            }
            // kt = Kf T, then K += T^T kt
            double kt[24][33] = {};
            for (usize r = 0; r < 24; ++r)
                for (usize m = 0; m < 8; ++m)
                    for (usize c = 0; c < 3; ++c) {
                        const double v = kf[r * 24 + 3 * m + c];
                        if (v != 0.0)
                            for (usize a = 0; a < 11; ++a) kt[r][3 * a + c] += v * phi[m][c][a];
                    }
            for (usize n = 0; n < 8; ++n)
                for (usize c = 0; c < 3; ++c) {
                    const double* row = kt[3 * n + c];
                    for (usize a = 0; a < 11; ++a) {
                        const double w = phi[n][c][a];
                        if (w == 0.0) continue;
                        double* o2 = &k[(3 * a + c) * 33];
                        for (usize j = 0; j < 33; ++j) o2[j] += w * row[j];
                    }
                }
        }
        out.cells.k.push_back(condense(k, coarse_modes(gram)));
        out.cells.origin.push_back(cc);
        start = end;
    }
    // assemble on the coarse unknowns
    std::vector<std::vector<std::pair<usize, u32>>> ids(out.cells.origin.size());
    std::vector<usize> eptr{0};
    std::vector<u32> conn;
    for (usize e = 0; e < out.cells.origin.size(); ++e) {
        const auto& o = out.cells.origin[e];
        for (usize n = 0; n < 8; ++n) {
            const std::array<u32, 3> q{o[0] + BITS[n][0], o[1] + BITS[n][1], o[2] + BITS[n][2]};
            for (usize c = 0; c < 3; ++c) {
                const u32 g = number[index(q, c, cd)];
                if (g != UINT32_MAX) {
                    ids[e].push_back({3 * n + c, g});
                    conn.push_back(g);
                }
            }
        }
        eptr.push_back(conn.size());
    }
    out.a = Csr::pattern(coarse.points.size(), eptr, conn);
    std::vector<double> ke;
    for (usize e = 0; e < ids.size(); ++e) {
        ke.clear();
        for (const auto& r : ids[e])
            for (const auto& c : ids[e]) ke.push_back(out.cells.k[e][r.first * 24 + c.first]);
        out.a.add_element(conn.data() + eptr[e], ids[e].size(), ke.data());
    }
    return out;
}

std::array<bool, 9> coarse_modes(const std::array<std::array<std::array<double, 11>, 11>, 3>& gram) {
    std::array<bool, 9> keep{};
    for (usize c = 0; c < 3; ++c) {
        const auto& g = gram[c];
        double l[11][11] = {};
        bool used[11] = {};
        for (usize j = 0; j < 11; ++j) {
            double d = g[j][j];
            for (usize k = 0; k < j; ++k)
                if (used[k]) d -= l[j][k] * l[j][k];
            if (!(d > 1e-6 * g[j][j])) continue;
            used[j] = true;
            d = std::sqrt(d);
            l[j][j] = d;
            for (usize i = j + 1; i < 11; ++i) {
                double v = g[i][j];
                for (usize k = 0; k < j; ++k)
                    if (used[k]) v -= l[i][k] * l[j][k];
                l[i][j] = v / d;
            }
            if (j >= 8) keep[3 * (j - 8) + c] = true;
        }
    }
    return keep;
}

std::array<double, 576> condense(std::vector<double>& k, const std::array<bool, 9>& keep) {
    for (usize m = 24; m < 33; ++m) {
        if (!keep[m - 24]) continue;
        const double pivot = k[m * 33 + m];
        double scale = 0.0;
        for (usize i = 0; i < 33; ++i) scale = std::max(scale, std::abs(k[i * 33 + i]));
        if (!(pivot > 1e-12 * scale)) continue;
        for (usize i = 0; i < 33; ++i) {
            const double f = k[i * 33 + m] / pivot;
            if (i == m || f == 0.0) continue;
            for (usize j = 0; j < 33; ++j) k[i * 33 + j] -= f * k[m * 33 + j];
        }
    }
    std::array<double, 576> out{};
    for (usize r = 0; r < 24; ++r)
        for (usize c = 0; c < 24; ++c) out[r * 24 + c] = k[r * 33 + c];
    return out;
}

Multigrid Multigrid::make(const Csr& a0, Level level, const CellMatrices* cells) {
    Multigrid mg;
    mg.a.push_back(a0);
    CellMatrices stored;
    bool have_stored = false;
    for (;;) {
        const Csr& a = mg.a.back();
        if (a.n <= COARSEST || mg.p.size() >= 30) break;
        auto c = coarsen(level, a);
        if (!c) break;
        if (c->coarse.points.size() * 10 > a.n * 9) break;
        mg.lines.push_back({Lines::make(level, 0, a), Lines::make(level, 1, a), Lines::make(level, 2, a)});
        Csr ac;
        if (!cells) {
            ac = c->p.galerkin(a);
        } else {
            ByCells bc = by_cells(level, have_stored ? stored : *cells, c->coarse, c->kept);
            ac = std::move(bc.a);
            stored = std::move(bc.cells);
            have_stored = true;
        }
        mg.a.push_back(std::move(ac));
        mg.p.push_back(std::move(c->p));
        level = std::move(c->coarse);
    }
    const Csr& a = mg.a.back();
    const auto xyz = level_xyz(level);
    const Symbolic sym = Symbolic::analyse(a, nested_dissection(a, xyz));
    mg.coarse = Cholesky::factor(sym, a);
    return mg;
}

void Multigrid::apply(const double* r, double* z) const { cycle(0, r, z); }

void Multigrid::cycle(usize l, const double* b, double* x) const {
    if (l + 1 == a.size()) {
        std::copy(b, b + a[l].n, x);
        coarse.solve(x);
        return;
    }
    const Csr& A = a[l];
    const Rect& P = p[l];
    std::fill(x, x + A.n, 0.0);
    for (const Lines& ln : lines[l]) ln.sweep(A, b, x, true);
    std::vector<double> r(A.n, 0.0);
    A.mul(x, r.data());
    for (usize i = 0; i < A.n; ++i) r[i] = b[i] - r[i];
    std::vector<double> rc(P.cols, 0.0);
    P.mul_transpose(r.data(), rc.data());
    std::vector<double> ec(P.cols, 0.0);
    cycle(l + 1, rc.data(), ec.data());
    P.mul(ec.data(), r.data());
    for (usize i = 0; i < A.n; ++i) x[i] += r[i];
    for (usize t = 3; t-- > 0;) lines[l][t].sweep(A, b, x, false);
}

// ------------------------------------------------------------------ pmultigrid.rs
PMultigrid PMultigrid::make(const Csr& a, Rect p, const std::vector<Vec3>& coarse) {
    PMultigrid out;
    const Csr ac = p.galerkin(a);
    const Symbolic sym = Symbolic::analyse(ac, nested_dissection(ac, coarse));
    out.coarse = Cholesky::factor(sym, ac);
    out.diag = a.diagonal();
    for (double d : out.diag)
        if (!(d > 0.0)) throw Error("the matrix has a diagonal entry that is not above 0");
    out.coarse_unknowns = p.cols;
    out.p = std::move(p);
    return out;
}

void PMultigrid::apply(const Csr& a, const double* r, double* z) const {
    const usize n = a.n;
    std::fill(z, z + n, 0.0);
    auto row = [&](const double* zz, usize i) {
        double s = r[i];
        for (usize k = a.ptr[i]; k < a.ptr[i + 1]; ++k) {
            const usize j = a.idx[k];
            if (j != i) s -= a.val[k] * zz[j];
        }
        return s / diag[i];
    };
    for (usize s = 0; s < SWEEPS; ++s)
        for (usize i = 0; i < n; ++i) z[i] = row(z, i);
    std::vector<double> az(n, 0.0);
    a.mul(z, az.data());
    std::vector<double> res(n);
    for (usize i = 0; i < n; ++i) res[i] = r[i] - az[i];
    std::vector<double> rc(p.cols, 0.0);
    p.mul_transpose(res.data(), rc.data());
    coarse.solve(rc.data());
    p.mul(rc.data(), az.data());
    for (usize i = 0; i < n; ++i) z[i] += az[i];
    for (usize s = 0; s < SWEEPS; ++s)
        for (usize i = n; i-- > 0;) z[i] = row(z, i);
}

std::pair<Rect, std::vector<Vec3>> prolongation(const std::vector<Vec3>& points, usize corners,
                                                const std::vector<std::array<u32, 2>>& edges,
                                                const std::vector<usize>& free) {
    std::vector<u32> coarse(3 * corners, UINT32_MAX);
    std::vector<Vec3> xyz;
    for (usize i : free)
        if (i < 3 * corners) {
            coarse[i] = (u32)xyz.size();
            xyz.push_back(points[i / 3]);
        }
    Rect P;
    P.ptr.reserve(free.size() + 1);
    P.ptr.push_back(0);
    for (usize i : free) {
        const usize v = i / 3, c = i % 3;
        if (v < corners) {
            P.idx.push_back(coarse[i]);
            P.val.push_back(1.0);
        } else {
            std::vector<u32> ends;
            for (u32 a : edges[v - corners]) {
                const u32 k = coarse[3 * (usize)a + c];
                if (k != UINT32_MAX) ends.push_back(k);
            }
            std::sort(ends.begin(), ends.end());
            for (u32 k : ends) {
                P.idx.push_back(k);
                P.val.push_back(0.5);
            }
        }
        P.ptr.push_back(P.idx.size());
    }
    P.rows = free.size();
    P.cols = xyz.size();
    return {std::move(P), std::move(xyz)};
}

// ------------------------------------------------------------------ solve.rs:92-250
Solver Solver::make(Csr a, Unknowns unknowns, std::optional<Method> method, Crossover crossover,
                    const CellMatrices* cells) {
    Solver s;
    auto choose = [&](const std::function<Symbolic()>& symbolic) -> std::optional<Symbolic> {
        if (method == Method::Direct) return symbolic();
        if (method == Method::Multigrid) return std::nullopt;
        if (!crossover.by_flops) {
            if (a.n <= crossover.unknowns) return symbolic();
            return std::nullopt;
        }
        Symbolic sym = symbolic();
        if (sym.entries() <= crossover.entries && (crossover.flat || sym.flops() <= crossover.flops * (double)a.n))
            return sym;
        return std::nullopt;
    };
    if (unknowns.kind == Unknowns::Quadratic) {
        auto sym = choose([&] { return Symbolic::analyse(a, nested_dissection(a, unknowns.points)); });
        if (sym) {
            s.chol = Cholesky::factor(*sym, a);
            s.inner = Direct;
        } else {
            s.pmg = PMultigrid::make(a, std::move(unknowns.p), unknowns.coarse);
            s.inner = PMG;
        }
        s.a = std::move(a);
        return s;
    }
    if (unknowns.kind == Unknowns::Points) {
        if (method == Method::Multigrid)
            throw Error("multigrid needs the box grid: a mesh of tetrahedra has the direct solver");
        const Symbolic sym = Symbolic::analyse(a, nested_dissection(a, unknowns.points));
        s.chol = Cholesky::factor(sym, a);
        s.inner = Direct;
        s.a = std::move(a);
        return s;
    }
    const Level& level = unknowns.level;
    auto sym = choose([&] { return Symbolic::analyse(a, nested_dissection(a, level_xyz(level))); });
    if (sym) {
        s.chol = Cholesky::factor(*sym, a);
        s.inner = Direct;
    } else {
        s.mg = Multigrid::make(a, level, cells);
        s.inner = MG;
    }
    s.a = std::move(a);
    return s;
}

Stats Solver::solve(const double* b, double* x, const Cancelled& cancelled) const {
    const usize n = a.n;
    if (inner == Direct) {
        std::copy(b, b + n, x);
        chol->solve(x);
        std::vector<double> ax(n, 0.0);
        a.mul(x, ax.data());
        double bn = 0.0, rn = 0.0;
        for (usize i = 0; i < n; ++i) bn += b[i] * b[i];
        for (usize i = 0; i < n; ++i) {
            const double d = b[i] - ax[i];
            rn += d * d;
        }
        bn = std::sqrt(bn);
        rn = std::sqrt(rn);
        return {0, bn > 0.0 ? rn / bn : rn};
    }
    Convergence c;
    if (inner == MG) {
        c = pcg(a, b, x, [&](const double* r, double* z) { mg->apply(r, z); }, tol, max_iter, cancelled);
    } else {
        c = pcg(a, b, x, [&](const double* r, double* z) { pmg->apply(a, r, z); }, tol, max_iter, cancelled);
    }
    return {c.iterations, c.residual};
}

}  // namespace orc
