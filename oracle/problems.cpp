// TEST INFRASTRUCTURE ONLY (the oracle). Structured box models assembled the way the reference
// callers assemble theirs:
//   hex8 heat:      heat.rs:755-877 (grid_model: number, pattern, conduction), heat.rs:424 (source
//                   V/8 per corner), heat.rs:503 (free), heat.rs:661-673 (submatrix, reduced_rhs,
//                   Unknowns::Grid)
//   tet4 heat:      heat.rs:880-968 (tet_model: tet_conduction), Unknowns::Points
//   hex elasticity: elastic.rs:287-563 (points numbered as heat's, unknown 3 point + c at
//                   elastic.rs:366-369, Hex elements, body force V/8 per corner at elastic.rs:402,
//                   free, Unknowns::Grid with the cells of multigrid.rs:67-70)
//   tet elasticity: elastic_tet.rs:173-313 (dofs 3 node + c, stiffness, body force -V/20 / V/5,
//                   face traction A/3 on the edge points, free, prolongation, Unknowns::Quadratic)
// The tet mesh of a box is the reference's extrusion (prism.rs:244-284) of a footprint of quads,
// each cut at its lowest-corner diagonal.
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

static std::vector<double> uniform_planes(int n, double len) {
    std::vector<double> p(n + 1);
    const double h = len / (double)n;
    for (int i = 0; i <= n; ++i) p[i] = h * (double)i;
    return p;
}

// heat.rs:1219-1242 with transient = None
std::vector<double> reduced_rhs(const Csr& a, const std::vector<double>& load, const std::vector<double>& t,
                                const std::vector<usize>& free) {
    std::vector<double> out(free.size());
    for (usize r = 0; r < free.size(); ++r) {
        const usize i = free[r];
        double v = load[i];
        for (usize e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
            const usize j = a.idx[e];
            if (!std::binary_search(free.begin(), free.end(), j)) v -= a.val[e] * t[j];
        }
        out[r] = v;
    }
    return out;
}

// The face of the box (0 x=0, 1 x=max, 2 y=0, 3 y=max, 4 z=0, 5 z=max) that grid point q is on.
static bool on_face(const std::array<u32, 3>& q, const std::array<u32, 3>& top, int face) {
    const int d = face / 2;
    return (face % 2 == 0) ? q[d] == 0 : q[d] == top[d];
}

Problem heat_hex8_box(const HeatBoxSpec& spec) {
    const BoxSpec& b = spec.box;
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    const std::array<std::vector<double>, 3> planes = {uniform_planes(nx, b.lx), uniform_planes(ny, b.ly),
                                                       uniform_planes(nz, b.lz)};
    const usize ncell = (usize)nx * ny * nz;
    if (spec.k.size() != ncell) throw Error("one conductivity for each cell");
    Problem pr;
    pr.comps = 1;
    pr.npe = 8;
    // heat.rs:1165-1208: first-seen numbering over the cells, corners in VTK order (grid.rs:201).
    std::vector<u32> number((usize)(nx + 1) * (ny + 1) * (nz + 1), UINT32_MAX);
    pr.cells.reserve(ncell * 8);
    static const int BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (usize cell = 0; cell < ncell; ++cell) {
        const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / ((usize)nx * ny));
        for (int c = 0; c < 8; ++c) {
            const std::array<u32, 3> p = {(u32)(i + BITS[c][0]), (u32)(j + BITS[c][1]), (u32)(k + BITS[c][2])};
            const usize slot = p[0] + (usize)(nx + 1) * (p[1] + (usize)(ny + 1) * p[2]);
            if (number[slot] == UINT32_MAX) {
                number[slot] = (u32)pr.ijk.size();
                pr.ijk.push_back(p);
            }
            pr.cells.push_back(number[slot]);
        }
    }
    const usize n = pr.ijk.size();
    pr.n_points = n;
    pr.points.resize(n);
    for (usize v = 0; v < n; ++v)
        for (int d = 0; d < 3; ++d) pr.points[v][d] = planes[d][pr.ijk[v][d]];
    std::vector<usize> eptr(ncell + 1);
    for (usize e = 0; e <= ncell; ++e) eptr[e] = 8 * e;
    pr.k = Csr::pattern(n, eptr, pr.cells);
    pr.load.assign(n, 0.0);
    for (usize cell = 0; cell < ncell; ++cell) {
        const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / ((usize)nx * ny));
        const Vec3 size = {planes[0][i + 1] - planes[0][i], planes[1][j + 1] - planes[1][j], planes[2][k + 1] - planes[2][k]};
        const double kc = spec.k[cell];
        const auto ke = conduction(size, {kc, kc, kc});
        pr.k.add_element(pr.cells.data() + 8 * cell, 8, ke.data());
        if (spec.source != 0.0) {
            const double vol = size[0] * size[1] * size[2];
            for (int c = 0; c < 8; ++c) pr.load[pr.cells[8 * cell + c]] += spec.source * vol / 8.0;
        }
    }
    const std::array<u32, 3> top = {(u32)nx, (u32)ny, (u32)nz};
    pr.is_fixed.assign(n, 0);
    pr.fixed_value.assign(n, 0.0);
    for (int f = 0; f < 6; ++f) {
        if (!spec.fixed[f]) continue;
        for (usize v = 0; v < n; ++v)
            if (on_face(pr.ijk[v], top, f)) {
                pr.is_fixed[v] = 1;
                pr.fixed_value[v] = spec.value[f];
            }
    }
    for (usize v = 0; v < n; ++v)
        if (!pr.is_fixed[v]) pr.free.push_back(v);
    pr.rhs = reduced_rhs(pr.k, pr.load, pr.fixed_value, pr.free);
    pr.unknowns.kind = Unknowns::Grid;
    pr.unknowns.level.planes = planes;
    for (usize i : pr.free) {
        pr.unknowns.level.points.push_back(pr.ijk[i]);
        pr.unknowns.level.comp.push_back(0);
    }
    return pr;
}

Problem elastic_hex8_box(const ElasticHexSpec& spec) {
    // the points and cells of the heat box (the same first-seen numbering)
    HeatBoxSpec hs;
    hs.box = spec.box;
    hs.k.assign((usize)spec.box.nx * spec.box.ny * spec.box.nz, 1.0);
    Problem grid = heat_hex8_box(hs);
    const BoxSpec& b = spec.box;
    const usize ncell = (usize)b.nx * b.ny * b.nz;
    if (spec.e.size() != ncell || spec.nu.size() != ncell) throw Error("one material for each cell");
    Problem pr;
    pr.comps = 3;
    pr.npe = 8;
    pr.n_points = grid.n_points;
    pr.points = grid.points;
    pr.ijk = grid.ijk;
    pr.cells = grid.cells;
    const usize n = 3 * pr.n_points;
    std::vector<usize> eptr(ncell + 1);
    std::vector<u32> conn;
    conn.reserve(24 * ncell);
    for (usize e = 0; e < ncell; ++e) {
        eptr[e] = 24 * e;
        for (int c = 0; c < 8; ++c)
            for (u32 q = 0; q < 3; ++q) conn.push_back(3 * pr.cells[8 * e + c] + q);
    }
    eptr[ncell] = 24 * ncell;
    pr.k = Csr::pattern(n, eptr, conn);
    pr.load.assign(n, 0.0);
    const auto& planes = grid.unknowns.level.planes;
    for (usize e = 0; e < ncell; ++e) {
        const int i = (int)(e % b.nx), j = (int)((e / b.nx) % b.ny), k = (int)(e / ((usize)b.nx * b.ny));
        const Vec3 size = {planes[0][i + 1] - planes[0][i], planes[1][j + 1] - planes[1][j], planes[2][k + 1] - planes[2][k]};
        const Stiffness d = Orthotropic::isotropic(spec.e[e], spec.nu[e]).stiffness();
        const auto ke = hex_stiffness(size, d);
        pr.k.add_element(conn.data() + 24 * e, 24, ke.data());
        pr.cell_matrices.origin.push_back({(u32)i, (u32)j, (u32)k});
        pr.cell_matrices.k.push_back(ke);
        const double vol = size[0] * size[1] * size[2];
        for (int c = 0; c < 8; ++c)
            for (int q = 0; q < 3; ++q) pr.load[conn[24 * e + 3 * c + q]] += spec.body[q] * vol / 8.0;
    }
    const std::array<u32, 3> top = {(u32)b.nx, (u32)b.ny, (u32)b.nz};
    pr.is_fixed.assign(n, 0);
    pr.fixed_value.assign(n, 0.0);
    for (int f = 0; f < 6; ++f) {
        if (!spec.clamped[f]) continue;
        for (usize v = 0; v < pr.n_points; ++v)
            if (on_face(pr.ijk[v], top, f))
                for (int q = 0; q < 3; ++q) pr.is_fixed[3 * v + q] = 1;
    }
    for (usize u = 0; u < n; ++u)
        if (!pr.is_fixed[u]) pr.free.push_back(u);
    pr.rhs = reduced_rhs(pr.k, pr.load, pr.fixed_value, pr.free);
    pr.unknowns.kind = Unknowns::Grid;
    pr.unknowns.level.planes = planes;
    for (usize u : pr.free) {
        pr.unknowns.level.points.push_back(pr.ijk[u / 3]);
        pr.unknowns.level.comp.push_back((std::uint8_t)(u % 3));
    }
    return pr;
}

// prism.rs:156-162
static double tet_volume(const std::array<Vec3, 4>& p) {
    auto d = [&](int k) -> Vec3 { return {p[k][0] - p[0][0], p[k][1] - p[0][1], p[k][2] - p[0][2]}; };
    const Vec3 a = d(1), b = d(2), c = d(3);
    return (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0])) /
           6.0;
}

void box_tets(const BoxSpec& box, std::vector<Vec3>& points, std::vector<std::array<u32, 3>>& ijk,
              std::vector<std::array<u32, 4>>& tets) {
    const int nx = box.nx, ny = box.ny, nz = box.nz;
    const auto px = uniform_planes(nx, box.lx), py = uniform_planes(ny, box.ly), pz = uniform_planes(nz, box.lz);
    const usize nv = (usize)(nx + 1) * (ny + 1);
    std::vector<u32> number(nv * (nz + 1), UINT32_MAX);
    points.clear();
    ijk.clear();
    tets.clear();
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const u32 a = (u32)(i + (nx + 1) * j), bq = a + 1, c = a + 1 + (u32)(nx + 1), dq = a + (u32)(nx + 1);
            const u32 tris[2][3] = {{a, bq, c}, {a, c, dq}};
            for (const auto& tri : tris)
                for (int k = 0; k < nz; ++k) {
                    std::array<std::uint64_t, 6> g{};
                    u32 node[6];
                    for (int cc = 0; cc < 6; ++cc) {
                        const usize plane = (usize)k + cc / 3;
                        g[cc] = plane * nv + tri[cc % 3];
                        u32& slot = number[g[cc]];
                        if (slot == UINT32_MAX) {
                            slot = (u32)points.size();
                            const u32 v2 = tri[cc % 3];
                            const u32 ii = v2 % (u32)(nx + 1), jj = v2 / (u32)(nx + 1);
                            points.push_back({px[ii], py[jj], pz[plane]});
                            ijk.push_back({ii, jj, (u32)plane});
                        }
                        node[cc] = slot;
                    }
                    for (const auto& local : split_prism(g)) {
                        std::array<u32, 4> tet = {node[local[0]], node[local[1]], node[local[2]], node[local[3]]};
                        const std::array<Vec3, 4> p = {points[tet[0]], points[tet[1]], points[tet[2]], points[tet[3]]};
                        if (tet_volume(p) < 0.0) std::swap(tet[2], tet[3]);
                        tets.push_back(tet);
                    }
                }
        }
}

static std::array<Vec3, 4> corners_of(const std::vector<Vec3>& points, const std::array<u32, 4>& t) {
    return {points[t[0]], points[t[1]], points[t[2]], points[t[3]]};
}

Problem heat_tet4_box(const HeatBoxSpec& spec) {
    Problem pr;
    pr.comps = 1;
    pr.npe = 4;
    box_tets(spec.box, pr.points, pr.ijk, pr.tets4);
    const usize n = pr.points.size(), nt = pr.tets4.size();
    if (spec.k.size() != nt) throw Error("one conductivity for each tetrahedron");
    pr.n_points = n;
    pr.corners = n;
    for (const auto& t : pr.tets4) pr.cells.insert(pr.cells.end(), t.begin(), t.end());
    std::vector<usize> eptr(nt + 1);
    for (usize e = 0; e <= nt; ++e) eptr[e] = 4 * e;
    pr.k = Csr::pattern(n, eptr, pr.cells);
    pr.load.assign(n, 0.0);
    for (usize e = 0; e < nt; ++e) {
        auto [g, v] = tet_gradients(corners_of(pr.points, pr.tets4[e]));
        const double kc = spec.k[e];
        const auto ke = tet_conduction(g, v, {kc, kc, kc});
        pr.k.add_element(pr.tets4[e].data(), 4, ke.data());
        if (spec.source != 0.0)
            for (int c = 0; c < 4; ++c) pr.load[pr.tets4[e][c]] += spec.source * v / 4.0;
    }
    const std::array<u32, 3> top = {(u32)spec.box.nx, (u32)spec.box.ny, (u32)spec.box.nz};
    pr.is_fixed.assign(n, 0);
    pr.fixed_value.assign(n, 0.0);
    for (int f = 0; f < 6; ++f) {
        if (!spec.fixed[f]) continue;
        for (usize v = 0; v < n; ++v)
            if (on_face(pr.ijk[v], top, f)) {
                pr.is_fixed[v] = 1;
                pr.fixed_value[v] = spec.value[f];
            }
    }
    for (usize v = 0; v < n; ++v)
        if (!pr.is_fixed[v]) pr.free.push_back(v);
    pr.rhs = reduced_rhs(pr.k, pr.load, pr.fixed_value, pr.free);
    pr.unknowns.kind = Unknowns::Points;
    for (usize i : pr.free) pr.unknowns.points.push_back(pr.points[i]);
    return pr;
}

Problem elastic_tet_box(const ElasticBoxSpec& spec) {
    if (spec.order != 4 && spec.order != 10) throw Error("the order is 4 or 10");
    Problem pr;
    pr.comps = 3;
    std::vector<Vec3> corner_points;
    std::vector<std::array<u32, 3>> corner_ijk;
    box_tets(spec.box, corner_points, corner_ijk, pr.tets4);
    const usize nv = corner_points.size(), nt = pr.tets4.size();
    if (spec.e.size() != nt || spec.nu.size() != nt) throw Error("one material for each tetrahedron");
    pr.corners = nv;
    const bool quad = spec.order == 10;
    const usize npe = quad ? 10 : 4;
    pr.npe = npe;
    Quadratic q;
    if (quad) {
        q = Quadratic::make(corner_points, pr.tets4);
        pr.points = q.points;
        pr.edges = q.edges;
        for (const auto& c : corner_ijk) pr.ijk.push_back({2 * c[0], 2 * c[1], 2 * c[2]});
        for (const auto& e : q.edges) {
            const auto& a = corner_ijk[e[0]];
            const auto& b = corner_ijk[e[1]];
            pr.ijk.push_back({a[0] + b[0], a[1] + b[1], a[2] + b[2]});
        }
        for (const auto& t : q.tets) pr.cells.insert(pr.cells.end(), t.begin(), t.end());
    } else {
        pr.points = corner_points;
        pr.ijk = corner_ijk;
        for (const auto& t : pr.tets4) pr.cells.insert(pr.cells.end(), t.begin(), t.end());
    }
    const usize n = pr.points.size();
    pr.n_points = n;
    // elastic_tet.rs:178-183: the unknown 3 node + component.
    std::vector<u32> dofs(nt * 3 * npe);
    for (usize e = 0; e < nt; ++e)
        for (usize k = 0; k < 3 * npe; ++k) dofs[e * 3 * npe + k] = 3 * pr.cells[e * npe + k / 3] + (u32)(k % 3);
    std::vector<usize> eptr(nt + 1);
    for (usize e = 0; e <= nt; ++e) eptr[e] = 3 * npe * e;
    pr.k = Csr::pattern(3 * n, eptr, dofs);
    pr.load.assign(3 * n, 0.0);
    for (usize e = 0; e < nt; ++e) {
        const Stiffness d = Orthotropic::isotropic(spec.e[e], spec.nu[e]).stiffness();
        const auto p = corners_of(corner_points, pr.tets4[e]);
        double volume;
        if (quad) {
            const Tet10 el = Tet10::make(p);
            volume = el.volume;
            pr.k.add_element(dofs.data() + e * 30, 30, el.stiffness(d).data());
        } else {
            auto [g, v] = tet_gradients(p);
            volume = v;
            pr.k.add_element(dofs.data() + e * 12, 12, tet4_stiffness(g, v, d).data());
        }
        if (spec.body[0] != 0.0 || spec.body[1] != 0.0 || spec.body[2] != 0.0) {
            for (usize i = 0; i < npe; ++i) {
                const double w = quad ? (i < 4 ? -volume / 20.0 : volume / 5.0) : volume / 4.0;
                const usize node = pr.cells[e * npe + i];
                for (int dd = 0; dd < 3; ++dd) pr.load[3 * node + dd] += spec.body[dd] * w;
            }
        }
    }
    // Face loads: each boundary triangle on a loaded box face gets A/3 on its 3 edge points
    // (tet10, elastic_tet.rs:253-266) or its 3 corners (tet4).
    const std::array<u32, 3> ctop = {(u32)spec.box.nx, (u32)spec.box.ny, (u32)spec.box.nz};
    static const int FACE[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int f = 0; f < 6; ++f) {
        const Vec3 t = spec.traction[f];
        if (t[0] == 0.0 && t[1] == 0.0 && t[2] == 0.0) continue;
        for (usize e = 0; e < nt; ++e) {
            const auto& tet = pr.tets4[e];
            for (const auto& fc : FACE) {
                const u32 a = tet[fc[0]], b = tet[fc[1]], c = tet[fc[2]];
                if (!(on_face(corner_ijk[a], ctop, f) && on_face(corner_ijk[b], ctop, f) && on_face(corner_ijk[c], ctop, f)))
                    continue;
                const Vec3 &p0 = corner_points[a], &p1 = corner_points[b], &p2 = corner_points[c];
                const Vec3 u = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
                const Vec3 w = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
                const Vec3 nn = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
                const double area = 0.5 * std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                u32 targets[3];
                if (quad) {
                    const auto& t10 = q.tets[e];
                    const int lv[3][2] = {{fc[0], fc[1]}, {fc[1], fc[2]}, {fc[2], fc[0]}};
                    for (int s = 0; s < 3; ++s) {
                        u32 mid = UINT32_MAX;
                        for (int k = 0; k < 6; ++k) {
                            const usize i0 = TET10_EDGES[k][0], i1 = TET10_EDGES[k][1];
                            if (((int)i0 == lv[s][0] && (int)i1 == lv[s][1]) || ((int)i0 == lv[s][1] && (int)i1 == lv[s][0]))
                                mid = t10[4 + k];
                        }
                        targets[s] = mid;
                    }
                } else {
                    targets[0] = a;
                    targets[1] = b;
                    targets[2] = c;
                }
                for (u32 node : targets)
                    for (int dd = 0; dd < 3; ++dd) pr.load[3 * (usize)node + dd] += t[dd] * area / 3.0;
            }
        }
    }
    const std::array<u32, 3> top = quad ? std::array<u32, 3>{2 * ctop[0], 2 * ctop[1], 2 * ctop[2]} : ctop;
    pr.is_fixed.assign(3 * n, 0);
    pr.fixed_value.assign(3 * n, 0.0);
    for (int f = 0; f < 6; ++f) {
        if (!spec.clamped[f]) continue;
        for (usize v = 0; v < n; ++v)
            if (on_face(pr.ijk[v], top, f))
                for (int dd = 0; dd < 3; ++dd) pr.is_fixed[3 * v + dd] = 1;
    }
    for (usize i = 0; i < 3 * n; ++i)
        if (!pr.is_fixed[i]) pr.free.push_back(i);
    pr.rhs = reduced_rhs(pr.k, pr.load, pr.fixed_value, pr.free);
    if (quad) {
        pr.unknowns.kind = Unknowns::Quadratic;
        for (usize i : pr.free) pr.unknowns.points.push_back(pr.points[i / 3]);
        auto pc = prolongation(pr.points, nv, q.edges, pr.free);
        pr.unknowns.p = std::move(pc.first);
        pr.unknowns.coarse = std::move(pc.second);
    } else {
        pr.unknowns.kind = Unknowns::Points;
        for (usize i : pr.free) pr.unknowns.points.push_back(pr.points[i / 3]);
    }
    return pr;
}

}  // namespace orc
