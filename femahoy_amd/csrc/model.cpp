// The structured box model: numbering, free unknowns, materials in grid layout, loads.
//
// Numbering rules kept alike with the reference so that the solution vector has the reference's
// layout:
//   hex8 points: first seen over the cells (x fastest), corners in VTK order
//                (crates/fem/src/heat.rs:1165-1208, grid.rs:201-213)
//   tet points:  the extrusion of a footprint of quads cut at their lowest corner, triangles in
//                footprint order and layers inside, points first seen (prism.rs:244-284), each
//                prism split by the smallest-index rule (prism.rs:131-154) and a tetrahedron of
//                negative volume reordered by swapping its corners 2 and 3 (prism.rs:278-280)
//   tet10 edges: first seen over the tetrahedra in the order of tet10.rs:38 EDGES, numbered after
//                the corners (tet10.rs:244-272)
//   unknowns:    3 node + component (elastic_tet.rs:178-183); the free ones increasing
//                (heat.rs:503, elastic_tet.rs:291)
#include "model.hpp"

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "hex_elastic.hpp"

namespace fem {

// The case of the split of a prism: m, the corner of the smallest global index, and a, whether
// the diagonal of the side opposite it starts at the corner next to m (prism.rs:131-154).
static void prism_case(const std::array<uint64_t, 6>& g, int& m, bool& a) {
    // Rotations that bring each corner to corner 0 (bottom 0 1 2, top 3 4 5 with 3 above 0).
    static const int ROT[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 2, 0, 4, 5, 3}, {2, 0, 1, 5, 3, 4},
                                  {3, 5, 4, 0, 2, 1}, {4, 3, 5, 1, 0, 2}, {5, 4, 3, 2, 1, 0}};
    m = 0;
    for (int k = 1; k < 6; ++k)
        if (g[k] < g[m]) m = k;
    const int* r = ROT[m];
    a = std::min(g[r[1]], g[r[5]]) < std::min(g[r[2]], g[r[4]]);
}

// The 3 tetrahedra (local corners) of the split of case (m, a).
static std::array<std::array<int, 4>, 3> prism_tets(int m, bool a) {
    static const int ROT[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 2, 0, 4, 5, 3}, {2, 0, 1, 5, 3, 4},
                                  {3, 5, 4, 0, 2, 1}, {4, 3, 5, 1, 0, 2}, {5, 4, 3, 2, 1, 0}};
    static const int CUT_A[3][4] = {{0, 1, 2, 5}, {0, 1, 5, 4}, {0, 4, 5, 3}};
    static const int CUT_B[3][4] = {{0, 1, 2, 4}, {0, 4, 2, 5}, {0, 4, 5, 3}};
    const int* r = ROT[m];
    const auto& cut = a ? CUT_A : CUT_B;
    std::array<std::array<int, 4>, 3> out{};
    for (int t = 0; t < 3; ++t)
        for (int v = 0; v < 4; ++v) out[t][v] = r[cut[t][v]];
    return out;
}

std::array<std::array<int, 4>, 3> prism_split(const std::array<uint64_t, 6>& g) {
    int m;
    bool a;
    prism_case(g, m, a);
    return prism_tets(m, a);
}

namespace {

const int VTK_BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int TET_EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
// Kuhn types by their axis order (the same table as the device code).
const int KUHN_ORDER[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

int kuhn_type(int a, int b) {
    for (int t = 0; t < 6; ++t)
        if (KUHN_ORDER[t][0] == a && KUHN_ORDER[t][1] == b) return t;
    return -1;
}

// The tetrahedra of a prism of the box grid, for each footprint triangle (0: corners (0,0),
// (1,0), (1,1) of the quad; 1: (0,0), (1,1), (0,1)) and each split case (m, a): their local corners
// with a tetrahedron of negative volume reordered by swapping its corners 2 and 3
// (prism.rs:278-280), and their Kuhn types. Every quantity follows from the corners' bits in the
// cell, so the table is built once from the same rules as a tetrahedron by tetrahedron pass.
struct PrismTable {
    int local[2][6][2][3][4];
    int type[2][6][2][3];
};

int kuhn_type(int a, int b);

const PrismTable& prism_table() {
    static const PrismTable table = [] {
        PrismTable tb{};
        const int QUAD[2][3][2] = {{{0, 0}, {1, 0}, {1, 1}}, {{0, 0}, {1, 1}, {0, 1}}};
        for (int tri = 0; tri < 2; ++tri)
            for (int m = 0; m < 6; ++m)
                for (int a = 0; a < 2; ++a) {
                    const auto tets = prism_tets(m, a != 0);
                    for (int t = 0; t < 3; ++t) {
                        int bits[4][3];
                        std::array<int, 4> loc = tets[t];
                        for (int v = 0; v < 4; ++v) {
                            const int cc = loc[v];
                            bits[v][0] = QUAD[tri][cc % 3][0];
                            bits[v][1] = QUAD[tri][cc % 3][1];
                            bits[v][2] = cc / 3;
                        }
                        long e[3][3];
                        for (int v = 0; v < 3; ++v)
                            for (int dd = 0; dd < 3; ++dd) e[v][dd] = bits[v + 1][dd] - bits[0][dd];
                        const long det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                                         e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                                         e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
                        if (det < 0) std::swap(loc[2], loc[3]);
                        for (int v = 0; v < 4; ++v) tb.local[tri][m][a][t][v] = loc[v];
                        // the chain order: vertices by the sum of their bits, the axis of each step
                        int sum[4], order[4] = {0, 1, 2, 3};
                        for (int v = 0; v < 4; ++v) sum[v] = bits[v][0] + bits[v][1] + bits[v][2];
                        std::sort(order, order + 4, [&](int x, int y) { return sum[x] < sum[y]; });
                        int axis[2] = {0, 0};
                        for (int st = 0; st < 2; ++st)
                            for (int dd = 0; dd < 3; ++dd)
                                if (bits[order[st + 1]][dd] != bits[order[st]][dd]) axis[st] = dd;
                        // -1 for a case whose tetrahedra are not Kuhn chains (no box ordering gives it)
                        tb.type[tri][m][a][t] = kuhn_type(axis[0], axis[1]);
                    }
                }
        return tb;
    }();
    return table;
}

bool on_face(const uint32_t q[3], const int top[3], int f) {
    const int d = f / 2;
    return (f % 2 == 0) ? q[d] == 0 : (int)q[d] == top[d];
}

// The loads of the box-grid elasticity caller on hex cells (elastic.rs:366-393 and 942-963):
// each cell puts body x V / 8 on each of its corners, in cell order; then each face with a
// traction t puts t A / 4 on each of the 4 points of each of its cell faces (A the cell face's
// area), the faces in the order x-, x+, y-, y+, z-, z+ and each face's cells in cell order.
void hex_loads(BoxModel& m) {
    const auto& in = m.in;
    const int nx = in.n[0], ny = in.n[1], nz = in.n[2];
    std::array<std::vector<double>, 3> pl;
    for (int d = 0; d < 3; ++d) {
        pl[d].resize((size_t)in.n[d] + 1);
        for (size_t i = 0; i < pl[d].size(); ++i) pl[d][i] = m.h[d] * (double)i;
    }
    auto gidx = [&](int i, int j, int k) { return (size_t)i + (size_t)m.gn[0] * ((size_t)j + (size_t)m.gn[1] * k); };
    const bool body = in.body[0] != 0.0 || in.body[1] != 0.0 || in.body[2] != 0.0;
    if (body)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double size[3] = {pl[0][i + 1] - pl[0][i], pl[1][j + 1] - pl[1][j], pl[2][k + 1] - pl[2][k]};
                    const double volume = size[0] * size[1] * size[2];
                    for (const auto& b : VTK_BITS)
                        for (int d = 0; d < 3; ++d)
                            m.load[3 * gidx(i + b[0], j + b[1], k + b[2]) + d] += in.body[d] * volume / 8.0;
                }
    for (int f = 0; f < 6; ++f) {
        const double* t = in.traction[f];
        if (t[0] == 0.0 && t[1] == 0.0 && t[2] == 0.0) continue;
        const int ax = f / 2, a = ax == 0 ? 1 : 0, b = ax == 2 ? 1 : 2;
        const int plane = (f % 2 == 0) ? 0 : in.n[ax];
        const int cell_plane = (f % 2 == 0) ? 0 : in.n[ax] - 1;
        int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        lo[ax] = cell_plane;
        hi[ax] = cell_plane + 1;
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    const double area = (pl[a][c[a] + 1] - pl[a][c[a]]) * (pl[b][c[b] + 1] - pl[b][c[b]]);
                    // the 4 points of the cell face: (0,0), (1,0), (1,1), (0,1) along its two axes
                    const int PTS[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
                    for (const auto& q : PTS) {
                        int p[3] = {i, j, k};
                        p[ax] = plane;
                        p[a] += q[0];
                        p[b] += q[1];
                        for (int d = 0; d < 3; ++d) m.load[3 * gidx(p[0], p[1], p[2]) + d] += t[d] * area / 4.0;
                    }
                }
    }
}

}  // namespace

BoxModel box_shape(const BoxInput& in) {
    for (int d = 0; d < 3; ++d) {
        if (in.n[d] < 1) throw Error("each axis needs 1 cell or more");
        if (!(in.len[d] > 0.0 && std::isfinite(in.len[d]))) throw Error("the box sizes must be finite numbers above 0");
    }
    BoxModel m;
    m.in = in;
    const size_t cells = (size_t)in.n[0] * in.n[1] * in.n[2];
    const bool tets = !hex_kind(in.kind);
    m.comps = (in.kind == Kind::ElasticTet4 || in.kind == Kind::ElasticTet10 || in.kind == Kind::ElasticHex8) ? 3 : 1;
    m.quadratic = in.kind == Kind::ElasticTet10;
    const int refine = m.quadratic ? 2 : 1;
    for (int d = 0; d < 3; ++d) {
        m.gn[d] = refine * in.n[d] + 1;
        m.h[d] = in.len[d] / (double)in.n[d];
    }
    for (int f = 0; f < 6; ++f)
        if (in.fixed[f]) m.fix |= 1u << f;
    const size_t need = tets ? 6 * cells : cells;
    if (in.coef0.size() != need)
        throw Error(tets ? "one material value for each tetrahedron"
                         : m.comps == 3 ? "one Young's modulus for each cell" : "one conductivity for each cell");
    if (m.comps == 3 && in.coef1.size() != need)
        throw Error(tets ? "one Poisson ratio for each tetrahedron" : "one Poisson ratio for each cell");
    // a node is free when it lies on no fixed face, that is, off the fixed faces along every axis
    size_t free_nodes = 1;
    for (int d = 0; d < 3; ++d) free_nodes *= (size_t)(m.gn[d] - (in.fixed[2 * d] ? 1 : 0) - (in.fixed[2 * d + 1] ? 1 : 0));
    m.free_count = free_nodes * m.comps;
    return m;
}

BoxModel build_model(const BoxInput& in) {
    BoxModel m = box_shape(in);
    const int nx = in.n[0], ny = in.n[1], nz = in.n[2];
    const size_t cells = (size_t)nx * ny * nz;
    const bool tets = !hex_kind(in.kind);

    const size_t gnodes = m.grid_nodes();
    auto gidx = [&](uint32_t i, uint32_t j, uint32_t k) { return i + (size_t)m.gn[0] * (j + (size_t)m.gn[1] * k); };

    // ---------------------------------------------------------------- numbering
    std::vector<uint32_t> gijk;  // grid point of each reference point, 3 per point
    // tet index (reference order) -> cell and Kuhn type
    std::vector<uint32_t>& tet_cell = m.tet_cell;
    std::vector<uint8_t>& tet_type = m.tet_type;
    if (!tets) {
        std::vector<uint32_t> number(gnodes, UINT32_MAX);
        for (size_t cell = 0; cell < cells; ++cell) {
            const uint32_t i = cell % nx, j = (cell / nx) % ny, k = cell / ((size_t)nx * ny);
            for (const auto& b : VTK_BITS) {
                const size_t gi = gidx(i + b[0], j + b[1], k + b[2]);
                if (number[gi] == UINT32_MAX) {
                    number[gi] = (uint32_t)m.point_grid.size();
                    m.point_grid.push_back((uint32_t)gi);
                }
            }
        }
    } else {
        const size_t nv = (size_t)(nx + 1) * (ny + 1);
        std::vector<uint32_t> number(nv * (nz + 1), UINT32_MAX);
        std::vector<std::array<uint32_t, 3>> corner_ijk;
        std::vector<std::array<uint32_t, 4>> tet_nodes;
        tet_nodes.reserve(6 * cells);
        tet_cell.reserve(6 * cells);
        tet_type.reserve(6 * cells);
        const PrismTable& pt = prism_table();
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const uint32_t a = i + (nx + 1) * j, b = a + 1, c = a + 1 + (nx + 1), d = a + (nx + 1);
                const uint32_t tris[2][3] = {{a, b, c}, {a, c, d}};
                for (int trix = 0; trix < 2; ++trix) {
                    const uint32_t* tri = tris[trix];
                    for (int k = 0; k < nz; ++k) {
                        std::array<uint64_t, 6> g{};
                        uint32_t node[6];
                        for (int cc = 0; cc < 6; ++cc) {
                            const uint64_t plane = (uint64_t)k + cc / 3;
                            g[cc] = plane * nv + tri[cc % 3];
                            uint32_t& slot = number[g[cc]];
                            if (slot == UINT32_MAX) {
                                slot = (uint32_t)corner_ijk.size();
                                const uint32_t v2 = tri[cc % 3];
                                corner_ijk.push_back({v2 % (uint32_t)(nx + 1), v2 / (uint32_t)(nx + 1), (uint32_t)plane});
                            }
                            node[cc] = slot;
                        }
                        int mcase;
                        bool acase;
                        prism_case(g, mcase, acase);
                        const uint32_t cell = (uint32_t)i + (uint32_t)nx * ((uint32_t)j + (uint32_t)ny * (uint32_t)k);
                        for (int t = 0; t < 3; ++t) {
                            const int* loc = pt.local[trix][mcase][acase][t];
                            if (pt.type[trix][mcase][acase][t] < 0) throw Error("the box split is not the Kuhn split");
                            if (m.quadratic)
                                tet_nodes.push_back({node[loc[0]], node[loc[1]], node[loc[2]], node[loc[3]]});
                            tet_cell.push_back(cell);
                            tet_type.push_back((uint8_t)pt.type[trix][mcase][acase][t]);
                        }
                    }
                }
            }
        const uint32_t scale = m.quadratic ? 2 : 1;
        for (const auto& c : corner_ijk) m.point_grid.push_back((uint32_t)gidx(scale * c[0], scale * c[1], scale * c[2]));
        if (m.quadratic) {
            std::vector<uint8_t> seen(gnodes, 0);
            for (const auto& t : tet_nodes)
                for (const auto& e : TET_EDGES) {
                    const auto& a = corner_ijk[t[e[0]]];
                    const auto& b = corner_ijk[t[e[1]]];
                    const size_t gi = gidx(a[0] + b[0], a[1] + b[1], a[2] + b[2]);
                    if (!seen[gi]) {
                        seen[gi] = 1;
                        m.point_grid.push_back((uint32_t)gi);
                    }
                }
        }
    }
    m.points = m.point_grid.size();
    if (m.points != gnodes) throw Error("the numbering does not cover the grid");

    // ---------------------------------------------------------------- materials
    if (in.kind == Kind::ElasticHex8) {
        m.cell_k = in.coef0;
        m.cell_nu = in.coef1;
        for (size_t e = 0; e < cells; ++e) (void)isotropic_stiffness(m.cell_k[e], m.cell_nu[e]);  // its texts
    } else if (!tets) {
        m.cell_k = in.coef0;
        for (double k : m.cell_k)
            if (!(k > 0.0 && std::isfinite(k))) throw Error("the conductivity must be a finite number above 0");
    } else {
        const int q = m.comps == 3 ? 2 : 1;
        m.tet_coef.assign((size_t)q * 6 * cells, 0.0);
        for (size_t t = 0; t < 6 * cells; ++t) {
            const size_t cell = tet_cell[t], type = tet_type[t];
            if (m.comps == 1) {
                const double k = in.coef0[t];
                if (!(k > 0.0 && std::isfinite(k))) throw Error("the conductivity must be a finite number above 0");
                m.tet_coef[type * cells + cell] = k;
            } else {
                const double e = in.coef0[t], nu = in.coef1[t];
                if (!(std::isfinite(e) && e > 0.0) || !std::isfinite(nu))
                    throw Error("the moduli must be above 0 and the Poisson ratios finite");
                if (!(nu > -1.0 && nu < 0.5))
                    throw Error(
                        "the Poisson ratios give no stable material (for an isotropic material, nu must be between "
                        "-1 and 0.5)");
                const double mu = e / (2.0 * (1.0 + nu));
                const double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
                m.tet_coef[(0 * 6 + type) * cells + cell] = lam;
                m.tet_coef[(1 * 6 + type) * cells + cell] = mu;
            }
        }
    }

    // ---------------------------------------------------------------- free unknowns
    const int top[3] = {m.gn[0] - 1, m.gn[1] - 1, m.gn[2] - 1};
    std::vector<uint8_t> node_fixed(gnodes, 0);
    m.fixed_value.assign(gnodes * m.comps, 0.0);
    for (size_t gi = 0; gi < gnodes; ++gi) {
        const uint32_t q[3] = {(uint32_t)(gi % m.gn[0]), (uint32_t)((gi / m.gn[0]) % m.gn[1]),
                               (uint32_t)(gi / ((size_t)m.gn[0] * m.gn[1]))};
        for (int f = 0; f < 6; ++f)
            if (in.fixed[f] && on_face(q, top, f)) {
                node_fixed[gi] = 1;
                if (m.comps == 1) m.fixed_value[gi] = in.value[f];
            }
    }
    m.free.reserve(m.free_count);
    m.free_grid.reserve(m.free_count);
    for (size_t p = 0; p < m.points; ++p) {
        const uint32_t gi = m.point_grid[p];
        if (node_fixed[gi]) continue;
        for (int c = 0; c < m.comps; ++c) {
            m.free.push_back((int64_t)(m.comps * p + c));
            m.free_grid.push_back((int64_t)(m.comps * (size_t)gi + c));
        }
    }
    if (m.free.size() != m.free_count) throw Error("internal: the free unknowns differ from the shape's count");

    // ---------------------------------------------------------------- loads
    m.load.assign(gnodes * m.comps, 0.0);
    const double hx = m.h[0], hy = m.h[1], hz = m.h[2];
    if (in.kind == Kind::HeatHex8 && in.source != 0.0) {
        const double v = hx * hy * hz;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (const auto& b : VTK_BITS) m.load[gidx(i + b[0], j + b[1], k + b[2])] += in.source * v / 8.0;
    }
    if (in.kind == Kind::HeatTet4 && in.source != 0.0) {
        const double v = hx * hy * hz / 6.0;
        // Each cell has 6 tetrahedra; corner 0 and corner 7 are in all 6, the others in 2.
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int t = 0; t < 6; ++t) {
                        int bits[3] = {0, 0, 0};
                        for (int vv = 0; vv < 4; ++vv) {
                            if (vv > 0) bits[KUHN_ORDER[t][vv - 1]] = 1;
                            m.load[gidx(i + bits[0], j + bits[1], k + bits[2])] += in.source * v / 4.0;
                        }
                    }
    }
    if (in.kind == Kind::ElasticHex8) hex_loads(m);
    if (m.comps == 3 && tets) {
        const bool body = in.body[0] != 0.0 || in.body[1] != 0.0 || in.body[2] != 0.0;
        const double v = hx * hy * hz / 6.0;
        const uint32_t s = m.quadratic ? 2 : 1;
        if (body)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i)
                        for (int t = 0; t < 6; ++t) {
                            uint32_t vx[4][3];
                            int bits[3] = {0, 0, 0};
                            for (int vv = 0; vv < 4; ++vv) {
                                if (vv > 0) bits[KUHN_ORDER[t][vv - 1]] = 1;
                                for (int dd = 0; dd < 3; ++dd) vx[vv][dd] = (uint32_t)(dd == 0 ? i : dd == 1 ? j : k) + bits[dd];
                            }
                            for (int vv = 0; vv < 4; ++vv) {
                                const size_t gi = gidx(s * vx[vv][0], s * vx[vv][1], s * vx[vv][2]);
                                const double w = m.quadratic ? -v / 20.0 : v / 4.0;
                                for (int dd = 0; dd < 3; ++dd) m.load[3 * gi + dd] += in.body[dd] * w;
                            }
                            if (m.quadratic)
                                for (const auto& e : TET_EDGES) {
                                    const size_t gi = gidx(vx[e[0]][0] + vx[e[1]][0], vx[e[0]][1] + vx[e[1]][1],
                                                           vx[e[0]][2] + vx[e[1]][2]);
                                    for (int dd = 0; dd < 3; ++dd) m.load[3 * gi + dd] += in.body[dd] * v / 5.0;
                                }
                        }
        // Face tractions: each boundary quad is 2 triangles cut along the diagonal from its lowest
        // corner; each triangle puts A/3 on its corners (tet4) or its edge points (tet10).
        for (int f = 0; f < 6; ++f) {
            const double* t = in.traction[f];
            if (t[0] == 0.0 && t[1] == 0.0 && t[2] == 0.0) continue;
            const int d = f / 2, u = d == 0 ? 1 : 0, w = d == 2 ? 1 : 2;
            const uint32_t plane = (f % 2 == 0) ? 0 : (uint32_t)in.n[d];
            const double area = 0.5 * m.h[u] * m.h[w];
            for (int b = 0; b < in.n[w]; ++b)
                for (int a = 0; a < in.n[u]; ++a) {
                    // corners of the quad in (u, w): p00, p10, p11, p01; triangles (p00,p10,p11), (p00,p11,p01)
                    auto corner = [&](int du, int dw) {
                        std::array<uint32_t, 3> q{};
                        q[d] = plane;
                        q[u] = a + du;
                        q[w] = b + dw;
                        return q;
                    };
                    const std::array<uint32_t, 3> tri[2][3] = {{corner(0, 0), corner(1, 0), corner(1, 1)},
                                                               {corner(0, 0), corner(1, 1), corner(0, 1)}};
                    for (const auto& tr : tri) {
                        for (int s2 = 0; s2 < 3; ++s2) {
                            size_t gi;
                            if (m.quadratic) {
                                const auto& p = tr[s2];
                                const auto& q2 = tr[(s2 + 1) % 3];
                                gi = gidx(p[0] + q2[0], p[1] + q2[1], p[2] + q2[2]);
                            } else {
                                gi = gidx(tr[s2][0], tr[s2][1], tr[s2][2]);
                            }
                            for (int dd = 0; dd < 3; ++dd) m.load[3 * gi + dd] += t[dd] * area / 3.0;
                        }
                    }
                }
        }
    }
    return m;
}

}  // namespace fem
