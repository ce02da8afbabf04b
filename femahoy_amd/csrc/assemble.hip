// Assembly on the device (include/fem_solver.h, fem_assemble): a mesh to the reduced matrix and
// right side that the reference's callers build on the CPU (crates/fem/src/sparse.rs:22-86
// Csr::pattern / add_element, sparse.rs:108-134 submatrix, heat.rs:1219-1242 reduced_rhs, with the
// element routines of heat.rs:972-1000 and 1246-1278 and tet10.rs:64-120).
//
// Host: the pattern, the free unknowns, the reduced pattern, and two gather maps: for every entry
// of the pattern the element-matrix slots that add_element adds into it, and for every unknown the
// element-load slots of its load, both in element order. Device: one thread per element-matrix row
// computes that row and the row's load; one thread per entry (per unknown) then sums its slots in
// the map's order starting from 0.0, which is the sequence of additions that the reference's
// element loop makes into that entry. This file is compiled without contraction of products into
// fused multiply-adds (-ffp-contract=off, and the pragma below), so the device's FP64 arithmetic
// is operation for operation a sequential CPU assembly in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fem_solver.h"
#include "../../include/fem_testing.h"
#include "common.hpp"
#include "hex_elastic.hpp"

#pragma clang fp contract(off)

namespace fem {
namespace {

constexpr int ABLOCK = 256;
constexpr uint32_t NONE32 = 0xffffffffu;

// Points per cell and unknowns per point of a mesh kind.
int points_per_cell(int kind) {
    return kind == FEM_MESH_HEAT_HEX8 || kind == FEM_MESH_ELASTIC_HEX8 ? 8 : kind == FEM_MESH_HEAT_TET4 ? 4 : 10;
}
int comps_of(int kind) { return kind == FEM_MESH_ELASTIC_TET10 || kind == FEM_MESH_ELASTIC_HEX8 ? 3 : 1; }

// ------------------------------------------------------------------ device: element rows
// Each thread computes row r of the matrix of one element into ke[e m^2 + r m ...] and the row's
// load into fe[e m + r].

// The barycentric gradients and the volume of a straight tetrahedron (heat.rs:983-1000).
__device__ void tet_gradients(const double p[4][3], double g[4][3], double& vol) {
    double e[3][3];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) e[k][d] = p[k + 1][d] - p[0][d];
    const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                       e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                       e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    // rows of the inverse: the cross products e1 x e2, e2 x e0, e0 x e1 over det
    const int A[3] = {1, 2, 0}, B[3] = {2, 0, 1};
    for (int r = 0; r < 3; ++r) {
        const double* a = e[A[r]];
        const double* b = e[B[r]];
        const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        for (int d = 0; d < 3; ++d) g[r + 1][d] = c[d] / det;
    }
    for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    vol = det / 6.0;
}

template <int NP>
__device__ void corners(const double* xyz, const uint32_t* conn, size_t e, double p[NP][3]) {
    for (int c = 0; c < NP; ++c) {
        const size_t q = conn[e * (size_t)NP + c];
        for (int d = 0; d < 3; ++d) p[c][d] = xyz[3 * q + d];
    }
}

// hex8 conduction of an axis-aligned box cell (heat.rs:1246-1278): sizes from corners 0 and 6.
__global__ void __launch_bounds__(ABLOCK) k_row_hex8(size_t cells, const double* xyz, const uint32_t* conn,
                                                     const double* k, double source, double* ke, double* fe) {
    const size_t t = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (t >= cells * 8) return;
    const size_t e = t / 8;
    const int p = (int)(t % 8);
    const uint32_t c0 = conn[8 * e], c6 = conn[8 * e + 6];
    double s[3];
    for (int d = 0; d < 3; ++d) s[d] = xyz[3 * (size_t)c6 + d] - xyz[3 * (size_t)c0 + d];
    const int BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const double kc = k[e];
    for (int q = 0; q < 8; ++q) {
        double v = 0.0;
        for (int d = 0; d < 3; ++d) {
            double term = kc;
            for (int a = 0; a < 3; ++a) {
                const bool same = BITS[p][a] == BITS[q][a];
                const double f = a == d ? (same ? 1.0 / s[a] : -1.0 / s[a]) : (same ? s[a] / 3.0 : s[a] / 6.0);
                term *= f;
            }
            v += term;
        }
        ke[e * 64 + (size_t)p * 8 + q] = v;
    }
    fe[e * 8 + p] = source * (s[0] * s[1] * s[2]) / 8.0;
}

// tet4 conduction (heat.rs:972-980).
__global__ void __launch_bounds__(ABLOCK) k_row_tet4(size_t cells, const double* xyz, const uint32_t* conn,
                                                     const double* k, double source, double* ke, double* fe) {
    const size_t t = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (t >= cells * 4) return;
    const size_t e = t / 4;
    const int i = (int)(t % 4);
    double p[4][3], g[4][3], vol;
    corners<4>(xyz, conn, e, p);
    tet_gradients(p, g, vol);
    const double kc = k[e];
    for (int j = 0; j < 4; ++j) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += kc * g[i][d] * g[j][d];
        ke[e * 16 + (size_t)i * 4 + j] = vol * s;
    }
    fe[e * 4 + i] = source * vol / 4.0;
}

// tet10 elasticity (tet10.rs:71-120): the 4-point rule of degree 2 (tet10.rs:44-48) over the
// quadratic shape gradients (tet10.rs:77-87); thread (e, i, a) computes row 3 i + a.
__global__ void __launch_bounds__(ABLOCK) k_row_tet10(size_t cells, const double* xyz, const uint32_t* conn,
                                                      const double* dmat, const double* body, double* ke,
                                                      double* fe) {
    const size_t t = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (t >= cells * 30) return;
    const size_t e = t / 30;
    const int row = (int)(t % 30), i = row / 3, a = row % 3;
    double p[4][3], g[4][3], vol;
    for (int c = 0; c < 4; ++c) {
        const size_t q = conn[e * 10 + c];
        for (int d = 0; d < 3; ++d) p[c][d] = xyz[3 * q + d];
    }
    tet_gradients(p, g, vol);
    double normal[3][3], shear[3];
    for (int r = 0; r < 9; ++r) normal[r / 3][r % 3] = dmat[e * 12 + r];
    for (int r = 0; r < 3; ++r) shear[r] = dmat[e * 12 + 9 + r];
    const int EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    const int SHEAR[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    const double qa = 0.5854101966249685, qb = 0.1381966011250105;
    const double w = vol / 4.0;
    double k[30];
    for (int c = 0; c < 30; ++c) k[c] = 0.0;
    for (int qp = 0; qp < 4; ++qp) {
        double l[4];
        for (int c = 0; c < 4; ++c) l[c] = c == qp ? qa : qb;
        double gr[10][3];
        for (int c = 0; c < 4; ++c)
            for (int d = 0; d < 3; ++d) gr[c][d] = (4.0 * l[c] - 1.0) * g[c][d];
        for (int m = 0; m < 6; ++m) {
            const int u = EDGES[m][0], v = EDGES[m][1];
            for (int d = 0; d < 3; ++d) gr[4 + m][d] = 4.0 * (l[u] * g[v][d] + l[v] * g[u][d]);
        }
        const double* gi = gr[i];
        for (int j = 0; j < 10; ++j) {
            const double* gj = gr[j];
            for (int b = 0; b < 3; ++b) {
                double v = gi[a] * gj[b] * normal[a][b];
                if (a == b) {
                    for (int s = 0; s < 3; ++s) {
                        const int ps = SHEAR[s][0], qs = SHEAR[s][1];
                        if (ps == a) v += gi[qs] * gj[qs] * shear[s];
                        else if (qs == a) v += gi[ps] * gj[ps] * shear[s];
                    }
                } else {
                    v += gi[b] * gj[a] * shear[3 - a - b];
                }
                k[3 * j + b] += w * v;
            }
        }
    }
    for (int c = 0; c < 30; ++c) ke[e * 900 + (size_t)row * 30 + c] = k[c];
    // the body force: -V/20 to each corner, V/5 to each edge node (elastic_tet.rs:202-230)
    const double wl = i < 4 ? -vol / 20.0 : vol / 5.0;
    fe[e * 30 + row] = body[a] * wl;
}

// hex8 elasticity (hex.rs:199-279, elastic.rs:366-393): row r of cell e is row r of its element,
// one of the distinct elements the host computed (table[which[e]]), and its load the body force
// times V/8 (V from corners 0 and 6), as the box-grid caller adds it.
__global__ void __launch_bounds__(ABLOCK) k_row_hex8e(size_t cells, const double* xyz, const uint32_t* conn,
                                                      const uint32_t* which, const double* table, const double* body,
                                                      double* ke, double* fe) {
    const size_t t = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (t >= cells * 24) return;
    const size_t e = t / 24;
    const int row = (int)(t % 24);
    const double* src = table + (size_t)which[e] * 576 + (size_t)row * 24;
    double* dst = ke + e * 576 + (size_t)row * 24;
    for (int c = 0; c < 24; ++c) dst[c] = src[c];
    const uint32_t c0 = conn[8 * e], c6 = conn[8 * e + 6];
    double s[3];
    for (int d = 0; d < 3; ++d) s[d] = xyz[3 * (size_t)c6 + d] - xyz[3 * (size_t)c0 + d];
    fe[e * 24 + row] = body[row % 3] * (s[0] * s[1] * s[2]) / 8.0;
}

// s = 0; s += src[slot] over the slots of item r in map order (the additions of the element loop).
__global__ void __launch_bounds__(ABLOCK) k_gather(size_t n, const uint64_t* ptr, const uint32_t* slot,
                                                   const double* src, const double* extra, double* out) {
    const size_t r = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (uint64_t c = ptr[r]; c < ptr[r + 1]; ++c) s += src[slot[c]];
    if (extra) s += extra[r];
    out[r] = s;
}

// The reduced values: entry k of the reduced matrix is entry from[k] of the whole pattern.
__global__ void __launch_bounds__(ABLOCK) k_pick(size_t n, const uint64_t* from, const double* val, double* out) {
    const size_t k = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (k < n) out[k] = val[from[k]];
}

// heat.rs:1219-1242 (steady): r_i = load_i - sum over the held columns j of row i, in column order,
// of a_ij t_j, for the free unknowns i.
__global__ void __launch_bounds__(ABLOCK) k_reduced_rhs(size_t nfree, const int64_t* free_list, const uint64_t* ptr,
                                                        const uint32_t* idx, const double* val, const uint8_t* fixed,
                                                        const double* value, const double* load, double* rhs) {
    const size_t f = (size_t)blockIdx.x * ABLOCK + threadIdx.x;
    if (f >= nfree) return;
    const size_t i = (size_t)free_list[f];
    double r = load[i];
    for (uint64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
        const uint32_t j = idx[e];
        if (fixed[j]) r -= val[e] * value[j];
    }
    rhs[f] = r;
}

unsigned grid_of(size_t n) { return (unsigned)((n + ABLOCK - 1) / ABLOCK); }

// ------------------------------------------------------------------ host
struct HostPlan {
    int kind = 0, npc = 0, comps = 1, m = 0;  // m: unknowns per element
    size_t units = 0, cells = 0;
    std::vector<uint32_t> eu;                 // the unknowns of each element (cells x m)
    std::vector<uint64_t> ptr;                // the whole pattern (sparse.rs:22-62)
    std::vector<uint32_t> idx;
    std::vector<int64_t> free_list;
    std::vector<uint64_t> rptr, rfrom;        // the reduced pattern and the entry each reduced entry is
    std::vector<uint32_t> ridx;
    std::vector<uint64_t> cptr;               // per entry of the pattern: its element-matrix slots
    std::vector<uint32_t> cslot;
    std::vector<uint64_t> lptr;               // per unknown: its element-load slots
    std::vector<uint32_t> lslot;
    std::vector<double> dmat;                 // tet10 elasticity: the 12 stiffness values of each cell
    std::vector<uint32_t> which;              // hex8 elasticity: each cell's element in `table`
    std::vector<double> table;                // hex8 elasticity: the distinct elements, 576 each
};

void validate(const fem_mesh& m) {
    if (m.kind < FEM_MESH_HEAT_HEX8 || m.kind > FEM_MESH_ELASTIC_HEX8) throw Error("unknown mesh kind");
    if (m.points == 0 || m.cells == 0) throw Error("the mesh needs points and cells");
    const bool elastic = m.kind == FEM_MESH_ELASTIC_TET10 || m.kind == FEM_MESH_ELASTIC_HEX8;
    if (!m.xyz || !m.conn || !m.fixed || !m.value) throw Error("the mesh needs xyz, conn, fixed and value");
    if (!m.coef0 && !(elastic && m.ortho)) throw Error("the mesh needs coef0 (or, for elasticity, ortho)");
    if (elastic && !m.coef1 && !m.ortho) throw Error("elasticity needs Poisson's ratio per cell (coef1)");
    if (m.ortho && !elastic) throw Error("orthotropic constants are for elasticity");
    const int npc = points_per_cell(m.kind);
    const uint64_t m2 = (uint64_t)(npc * comps_of(m.kind)) * (uint64_t)(npc * comps_of(m.kind));
    if (m.cells * m2 >= (uint64_t)NONE32) throw Error("the mesh has too many cells for 32-bit element slots");
    if (m.points * comps_of(m.kind) >= (uint64_t)NONE32) throw Error("the mesh has too many points");
    for (uint64_t c = 0; c < m.cells * npc; ++c)
        if (m.conn[c] >= m.points) throw Error("a cell names a point that the mesh does not have");
}

// The counting sort of the slots 0 .. targets.size() - 1 by their target, stable: for every target
// its slots in increasing order (the element order).
void group_by_target(const std::vector<uint32_t>& target, size_t ntargets, std::vector<uint64_t>& ptr,
                     std::vector<uint32_t>& slot) {
    ptr.assign(ntargets + 1, 0);
    for (uint32_t t : target) ++ptr[(size_t)t + 1];
    for (size_t i = 0; i < ntargets; ++i) ptr[i + 1] += ptr[i];
    slot.resize(target.size());
    std::vector<uint64_t> next(ptr.begin(), ptr.end() - 1);
    for (size_t c = 0; c < target.size(); ++c) slot[next[target[c]]++] = (uint32_t)c;
}

void run_parallel(size_t n, const std::function<void(size_t, size_t)>& f) {
    const size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    if (n < 4096 || nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t) th.emplace_back(f, n * t / nt, n * (t + 1) / nt);
    for (auto& x : th) x.join();
}

// The host part: the element unknowns, the pattern, the free list and reduced pattern; with maps,
// also the two gather maps and the material stiffness.
void plan(const fem_mesh& mesh, HostPlan& P, bool maps) {
    validate(mesh);
    P.kind = mesh.kind;
    P.npc = points_per_cell(mesh.kind);
    P.comps = comps_of(mesh.kind);
    P.m = P.npc * P.comps;
    P.cells = mesh.cells;
    P.units = mesh.points * P.comps;
    const int m = P.m;
    // elastic_tet.rs:178-183: unknown 3 point + component, in the order of the element's points
    P.eu.resize(P.cells * m);
    for (size_t e = 0; e < P.cells; ++e)
        for (int k = 0; k < m; ++k) P.eu[e * m + k] = P.comps * mesh.conn[e * P.npc + k / P.comps] + (uint32_t)(k % P.comps);
    // sparse.rs:22-62: each unknown couples to itself and to every unknown of every element that
    // holds it, the row sorted
    const size_t n = P.units;
    std::vector<uint64_t> cnt(n + 1, 0);
    for (uint32_t u : P.eu) ++cnt[(size_t)u + 1];
    for (size_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint32_t> of_unit(cnt[n]);
    {
        std::vector<uint64_t> nx(cnt.begin(), cnt.end() - 1);
        for (size_t e = 0; e < P.cells; ++e)
            for (int k = 0; k < m; ++k) of_unit[nx[P.eu[e * m + k]]++] = (uint32_t)e;
    }
    std::vector<std::vector<uint32_t>> rows(n);
    run_parallel(n, [&](size_t lo, size_t hi) {
        std::vector<uint32_t> mark(n, NONE32);  // per thread: the last row that saw each unknown
        for (size_t i = lo; i < hi; ++i) {
            auto& row = rows[i];
            mark[i] = (uint32_t)i;
            row.push_back((uint32_t)i);
            for (uint64_t c = cnt[i]; c < cnt[i + 1]; ++c) {
                const uint32_t* el = &P.eu[(size_t)of_unit[c] * m];
                for (int b = 0; b < m; ++b)
                    if (mark[el[b]] != (uint32_t)i) {
                        mark[el[b]] = (uint32_t)i;
                        row.push_back(el[b]);
                    }
            }
            std::sort(row.begin(), row.end());
        }
    });
    P.ptr.assign(n + 1, 0);
    for (size_t i = 0; i < n; ++i) P.ptr[i + 1] = P.ptr[i] + rows[i].size();
    P.idx.resize(P.ptr[n]);
    for (size_t i = 0; i < n; ++i) {
        std::copy(rows[i].begin(), rows[i].end(), P.idx.begin() + (ptrdiff_t)P.ptr[i]);
        std::vector<uint32_t>().swap(rows[i]);
    }
    // the free unknowns, increasing (heat.rs:503, elastic_tet.rs:291), and submatrix (sparse.rs:108-134)
    std::vector<uint32_t> renum(n, NONE32);
    P.free_list.clear();
    for (size_t u = 0; u < n; ++u)
        if (!mesh.fixed[u]) {
            renum[u] = (uint32_t)P.free_list.size();
            P.free_list.push_back((int64_t)u);
        }
    P.rptr.assign(P.free_list.size() + 1, 0);
    P.ridx.clear();
    P.rfrom.clear();
    for (size_t f = 0; f < P.free_list.size(); ++f) {
        const size_t i = (size_t)P.free_list[f];
        for (uint64_t e = P.ptr[i]; e < P.ptr[i + 1]; ++e) {
            const uint32_t j = renum[P.idx[e]];
            if (j != NONE32) {
                P.ridx.push_back(j);
                P.rfrom.push_back(e);
            }
        }
        P.rptr[f + 1] = P.ridx.size();
    }
    if (!maps) return;
    // the entry that each element-matrix slot adds into (add_element, sparse.rs:64-86)
    std::vector<uint32_t> target(P.cells * (size_t)m * m);
    run_parallel(P.cells, [&](size_t lo, size_t hi) {
        for (size_t e = lo; e < hi; ++e) {
            const uint32_t* el = &P.eu[e * m];
            for (int a = 0; a < m; ++a) {
                const uint32_t* r0 = P.idx.data() + P.ptr[el[a]];
                const uint32_t* r1 = P.idx.data() + P.ptr[el[a] + 1];
                for (int b = 0; b < m; ++b) {
                    const uint32_t* at = std::lower_bound(r0, r1, el[b]);
                    target[(e * m + a) * m + b] = (uint32_t)(P.ptr[el[a]] + (uint64_t)(at - r0));
                }
            }
        }
    });
    if (P.idx.size() >= (size_t)NONE32) throw Error("the pattern has too many entries for 32-bit entry numbers");
    group_by_target(target, P.idx.size(), P.cptr, P.cslot);
    // the unknown that each element-load slot adds into (the callers' load loops, element order)
    group_by_target(P.eu, n, P.lptr, P.lslot);
    if (P.kind == FEM_MESH_ELASTIC_HEX8) {
        // the element of each distinct (sizes, material): sizes from corners 0 and 6 (grid.rs:195-198)
        std::map<std::array<uint64_t, 12>, uint32_t> seen;
        P.which.resize(P.cells);
        P.table.clear();
        auto bits = [](double v) {
            uint64_t b;
            std::memcpy(&b, &v, sizeof b);
            return b;
        };
        for (size_t e = 0; e < P.cells; ++e) {
            const uint32_t c0 = mesh.conn[8 * e], c6 = mesh.conn[8 * e + 6];
            std::array<double, 3> size;
            for (int d = 0; d < 3; ++d) size[d] = mesh.xyz[3 * (size_t)c6 + d] - mesh.xyz[3 * (size_t)c0 + d];
            std::array<uint64_t, 12> key = {bits(size[0]), bits(size[1]), bits(size[2])};
            if (mesh.ortho)
                for (int q = 0; q < 9; ++q) key[3 + q] = bits(mesh.ortho[9 * e + q]);
            else
                key[3] = bits(mesh.coef0[e]), key[4] = bits(mesh.coef1[e]);
            auto it = seen.find(key);
            if (it == seen.end()) {
                if (!(size[0] > 0.0 && size[1] > 0.0 && size[2] > 0.0))
                    throw Error("a hex cell needs corner 6 above corner 0 along every axis");
                it = seen.emplace(key, (uint32_t)(P.table.size() / 576)).first;
                const auto k = hex_element(size, mesh.ortho ? orthotropic_stiffness(mesh.ortho + 9 * e)
                                                            : isotropic_stiffness(mesh.coef0[e], mesh.coef1[e]));
                P.table.insert(P.table.end(), k.begin(), k.end());
            }
            P.which[e] = it->second;
        }
    }
    if (P.kind == FEM_MESH_ELASTIC_TET10) {
        P.dmat.resize(P.cells * 12);
        for (size_t e = 0; e < P.cells; ++e) {
            const auto d = mesh.ortho ? orthotropic_stiffness(mesh.ortho + 9 * e)
                                      : isotropic_stiffness(mesh.coef0[e], mesh.coef1[e]);
            std::copy(d.begin(), d.end(), P.dmat.begin() + (ptrdiff_t)(e * 12));
        }
    }
}

template <class T>
DArray<T> up(const std::vector<T>& v, hipStream_t s) {
    DArray<T> d(std::max<size_t>(v.size(), 1));
    if (!v.empty()) d.upload(v.data(), v.size(), s);
    return d;
}
template <class T>
DArray<T> up(const T* p, size_t n, hipStream_t s) {
    DArray<T> d(std::max<size_t>(n, 1));
    if (n) d.upload(p, n, s);
    return d;
}

}  // namespace

class Assembly {
   public:
    Assembly(const fem_mesh& mesh, int device) {
        // the host part first: a bad mesh or material is refused before the device is touched
        const auto t0 = std::chrono::steady_clock::now();
        HostPlan P;
        plan(mesh, P, true);
        const auto t1 = std::chrono::steady_clock::now();
        st_.host_seconds = std::chrono::duration<double>(t1 - t0).count();
        FEM_HIP(hipSetDevice(device));
        hipStream_t s;
        FEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        FEM_HIP(hipEventCreate(&e0));
        FEM_HIP(hipEventCreate(&e1));
        try {
            run(mesh, P, s, e0, e1);
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            (void)hipStreamDestroy(s);
            throw;
        }
        FEM_HIP(hipEventDestroy(e0));
        FEM_HIP(hipEventDestroy(e1));
        FEM_HIP(hipStreamDestroy(s));
        st_.free = P.free_list.size();
        st_.entries = P.ridx.size();
        st_.full_entries = P.idx.size();
        free_list_ = std::move(P.free_list);
    }

    const fem_assembly_stats& stats() const { return st_; }

    void download(uint64_t* ptr, uint32_t* idx, double* val, double* rhs, int64_t* free_list) const {
        const size_t nf = (size_t)st_.free, ne = (size_t)st_.entries;
        if (ptr) FEM_HIP(hipMemcpy(ptr, rptr_.data(), (nf + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (idx && ne) FEM_HIP(hipMemcpy(idx, ridx_.data(), ne * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (val && ne) FEM_HIP(hipMemcpy(val, rval_.data(), ne * sizeof(double), hipMemcpyDeviceToHost));
        if (rhs && nf) FEM_HIP(hipMemcpy(rhs, rhs_.data(), nf * sizeof(double), hipMemcpyDeviceToHost));
        if (free_list && nf) std::memcpy(free_list, free_list_.data(), nf * sizeof(int64_t));
    }
    void device(const uint64_t** ptr, const uint32_t** idx, const double** val, const double** rhs) const {
        if (ptr) *ptr = rptr_.data();
        if (idx) *idx = ridx_.data();
        if (val) *val = rval_.data();
        if (rhs) *rhs = rhs_.data();
    }

   private:
    void run(const fem_mesh& mesh, const HostPlan& P, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
        const size_t n = P.units, m = (size_t)P.m, nc = P.cells;
        const int npc = P.npc;
        // the mesh and the maps to the device (the uploads are outside the device time)
        DArray<double> xyz = up(mesh.xyz, mesh.points * 3, s);
        DArray<uint32_t> conn = up(mesh.conn, nc * npc, s);
        // the D matrices of tet10, the conductivities of heat; hex8 elasticity reads its element table
        DArray<double> coef = P.kind == FEM_MESH_ELASTIC_TET10 ? up(P.dmat, s)
                              : P.kind == FEM_MESH_ELASTIC_HEX8 ? DArray<double>(1) : up(mesh.coef0, nc, s);
        DArray<double> body = up(mesh.body, 3, s);
        DArray<uint8_t> fixed = up(mesh.fixed, n, s);
        DArray<double> value = up(mesh.value, n, s);
        DArray<double> extra;
        if (mesh.load) extra = up(mesh.load, n, s);
        DArray<uint64_t> ptr = up(P.ptr, s), cptr = up(P.cptr, s), lptr = up(P.lptr, s), rfrom = up(P.rfrom, s);
        DArray<uint32_t> idx = up(P.idx, s), cslot = up(P.cslot, s), lslot = up(P.lslot, s);
        DArray<int64_t> fl = up(P.free_list, s);
        rptr_ = up(P.rptr, s);
        ridx_ = up(P.ridx, s);
        DArray<double> ke(nc * m * m), fe(nc * m), val(std::max<size_t>(P.idx.size(), 1)), load(n);
        rval_.resize(std::max<size_t>(P.ridx.size(), 1));
        rhs_.resize(std::max<size_t>(P.free_list.size(), 1));
        DArray<uint32_t> which;
        DArray<double> table;
        if (P.kind == FEM_MESH_ELASTIC_HEX8) {
            which = up(P.which, s);
            table = up(P.table, s);
        }
        FEM_HIP(hipEventRecord(e0, s));
        const size_t rows = nc * m;
        switch (P.kind) {
            case FEM_MESH_HEAT_HEX8:
                hipLaunchKernelGGL(k_row_hex8, dim3(grid_of(rows)), dim3(ABLOCK), 0, s, nc, (const double*)xyz.data(),
                                   (const uint32_t*)conn.data(), (const double*)coef.data(), mesh.source, ke.data(),
                                   fe.data());
                break;
            case FEM_MESH_HEAT_TET4:
                hipLaunchKernelGGL(k_row_tet4, dim3(grid_of(rows)), dim3(ABLOCK), 0, s, nc, (const double*)xyz.data(),
                                   (const uint32_t*)conn.data(), (const double*)coef.data(), mesh.source, ke.data(),
                                   fe.data());
                break;
            case FEM_MESH_ELASTIC_HEX8:
                hipLaunchKernelGGL(k_row_hex8e, dim3(grid_of(rows)), dim3(ABLOCK), 0, s, nc, (const double*)xyz.data(),
                                   (const uint32_t*)conn.data(), (const uint32_t*)which.data(),
                                   (const double*)table.data(), (const double*)body.data(), ke.data(), fe.data());
                break;
            default:
                hipLaunchKernelGGL(k_row_tet10, dim3(grid_of(rows)), dim3(ABLOCK), 0, s, nc, (const double*)xyz.data(),
                                   (const uint32_t*)conn.data(), (const double*)coef.data(), (const double*)body.data(),
                                   ke.data(), fe.data());
        }
        FEM_HIP(hipGetLastError());
        const size_t nnz = P.idx.size();
        hipLaunchKernelGGL(k_gather, dim3(grid_of(nnz)), dim3(ABLOCK), 0, s, nnz, (const uint64_t*)cptr.data(),
                           (const uint32_t*)cslot.data(), (const double*)ke.data(), (const double*)nullptr, val.data());
        hipLaunchKernelGGL(k_gather, dim3(grid_of(n)), dim3(ABLOCK), 0, s, n, (const uint64_t*)lptr.data(),
                           (const uint32_t*)lslot.data(), (const double*)fe.data(),
                           (const double*)(mesh.load ? extra.data() : nullptr), load.data());
        const size_t nr = P.ridx.size(), nf = P.free_list.size();
        if (nr)
            hipLaunchKernelGGL(k_pick, dim3(grid_of(nr)), dim3(ABLOCK), 0, s, nr, (const uint64_t*)rfrom.data(),
                               (const double*)val.data(), rval_.data());
        if (nf)
            hipLaunchKernelGGL(k_reduced_rhs, dim3(grid_of(nf)), dim3(ABLOCK), 0, s, nf, (const int64_t*)fl.data(),
                               (const uint64_t*)ptr.data(), (const uint32_t*)idx.data(), (const double*)val.data(),
                               (const uint8_t*)fixed.data(), (const double*)value.data(), (const double*)load.data(),
                               rhs_.data());
        FEM_HIP(hipGetLastError());
        FEM_HIP(hipEventRecord(e1, s));
        FEM_HIP(hipStreamSynchronize(s));  // the temporaries above are freed when this returns
        float ms = 0.f;
        FEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        st_.device_seconds = ms * 1e-3;
    }

    fem_assembly_stats st_{};
    DArray<uint64_t> rptr_;
    DArray<uint32_t> ridx_;
    DArray<double> rval_, rhs_;
    std::vector<int64_t> free_list_;
};

}  // namespace fem

struct fem_assembly {
    std::unique_ptr<fem::Assembly> impl;
};

namespace {
void put_msg(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}
template <class F>
int guard_call(char* err, size_t errlen, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        put_msg(err, errlen, e.what());
        return 1;
    }
}
}  // namespace

extern "C" {

int fem_assemble(const fem_mesh* mesh, int device, fem_assembly** out, char* err, size_t errlen) {
    return guard_call(err, errlen, [&] {
        if (!mesh || !out) throw fem::Error("fem_assemble needs a mesh and an output handle");
        *out = nullptr;
        auto a = std::make_unique<fem_assembly>();
        a->impl = std::make_unique<fem::Assembly>(*mesh, device);
        *out = a.release();
    });
}

void fem_assembly_stats_of(const fem_assembly* a, fem_assembly_stats* out) {
    if (a && out) *out = a->impl->stats();
}

int fem_assembly_download(const fem_assembly* a, uint64_t* ptr, uint32_t* idx, double* val, double* rhs,
                          int64_t* free_list, char* err, size_t errlen) {
    return guard_call(err, errlen, [&] {
        if (!a) throw fem::Error("no assembly");
        a->impl->download(ptr, idx, val, rhs, free_list);
    });
}

void fem_assembly_device(const fem_assembly* a, const uint64_t** ptr, const uint32_t** idx, const double** val,
                         const double** rhs) {
    if (a) a->impl->device(ptr, idx, val, rhs);
}

void fem_assembly_free(fem_assembly* a) { delete a; }

int fem_test_assembly_pattern(const fem_mesh* mesh, uint64_t* free_count, uint64_t* entries, uint64_t* ptr,
                              uint32_t* idx, int64_t* free_list, char* err, size_t errlen) {
    return guard_call(err, errlen, [&] {
        if (!mesh) throw fem::Error("no mesh");
        fem::HostPlan P;
        fem::plan(*mesh, P, false);
        if (free_count) *free_count = P.free_list.size();
        if (entries) *entries = P.ridx.size();
        if (ptr) std::copy(P.rptr.begin(), P.rptr.end(), ptr);
        if (idx) std::copy(P.ridx.begin(), P.ridx.end(), idx);
        if (free_list) std::copy(P.free_list.begin(), P.free_list.end(), free_list);
    });
}

}  // extern "C"
