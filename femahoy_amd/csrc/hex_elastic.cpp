// Linear elasticity on the box grid's hex cells: the condensed incompatible-mode element and the
// reduced system of a box (crates/fem/src/hex.rs:146-279, crates/fem/src/elastic.rs:287-547).
// No FMA contraction in this file: each sum is the reference's sequence of FP64 operations.
#include "hex_elastic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <tuple>

#include "common.hpp"

#pragma clang fp contract(off)

namespace fem {

namespace {

// The corners of a cell by their bits along x, y, z, in the order of a VTK hexahedron (hex.rs:34).
constexpr int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// The corner of each bit pattern x + 2 y + 4 z.
constexpr int CORNER_OF_BITS[8] = {0, 1, 3, 2, 4, 5, 7, 6};
// The two axes of each engineering shear strain: yz, xz, xy (hex.rs:126).
constexpr int SHEAR_AXES[3][2] = {{1, 2}, {0, 2}, {0, 1}};

// Gradients of the 8 corner functions and the 3 mode functions 1 - xi_d^2 at the local point xi
// of a box with sides `size` (hex.rs:130-146).
void gradients11(const std::array<double, 3>& size, const double xi[3], double g[11][3]) {
    for (int n = 0; n < 8; ++n) {
        double sg[3], f[3];
        for (int d = 0; d < 3; ++d) {
            sg[d] = CORNER[n][d] == 1 ? 1.0 : -1.0;
            f[d] = 0.5 * (1.0 + sg[d] * xi[d]);
        }
        for (int d = 0; d < 3; ++d) g[n][d] = sg[d] / size[d] * f[(d + 1) % 3] * f[(d + 2) % 3];
    }
    for (int m = 8; m < 11; ++m)
        for (int d = 0; d < 3; ++d) g[m][d] = 0.0;
    for (int d = 0; d < 3; ++d) g[8 + d][d] = -4.0 * xi[d] / size[d];
}

// The lower Cholesky factor of a 9 x 9 matrix and the solve with it (hex.rs:364-398).
void factor9(const double a[9][9], double l[9][9]) {
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) l[r][c] = 0.0;
    for (int j = 0; j < 9; ++j) {
        double d = a[j][j];
        for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
        d = std::sqrt(d);
        l[j][j] = d;
        for (int i = j + 1; i < 9; ++i) {
            double v = a[i][j];
            for (int k = 0; k < j; ++k) v -= l[i][k] * l[j][k];
            l[i][j] = v / d;
        }
    }
}
void solve9(const double l[9][9], double b[9]) {
    for (int i = 0; i < 9; ++i) {
        for (int k = 0; k < i; ++k) b[i] -= l[i][k] * b[k];
        b[i] /= l[i][i];
    }
    for (int i = 8; i >= 0; --i) {
        for (int k = i + 1; k < 9; ++k) b[i] -= l[k][i] * b[k];
        b[i] /= l[i][i];
    }
}

void parallel_rows(size_t n, const std::function<void(size_t, size_t)>& f) {
    const size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
    if (n < 8192 || nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t) th.emplace_back(f, n * t / nt, n * (t + 1) / nt);
    for (auto& x : th) x.join();
}

uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

}  // namespace

std::array<double, 12> isotropic_stiffness(double e, double nu) {
    const double g = e / (2.0 * (1.0 + nu));
    const double m[9] = {e, e, e, nu, nu, nu, g, g, g};
    return orthotropic_stiffness(m);
}

std::array<double, 12> orthotropic_stiffness(const double m[9]) {
    const double ex = m[0], ey = m[1], ez = m[2], nxy = m[3], nxz = m[4], nyz = m[5];
    bool moduli = true;
    for (int i : {0, 1, 2, 6, 7, 8}) moduli = moduli && std::isfinite(m[i]) && m[i] > 0.0;
    if (!moduli || !std::isfinite(nxy) || !std::isfinite(nxz) || !std::isfinite(nyz))
        throw Error("the moduli must be above 0 and the Poisson ratios finite");
    const double s[3][3] = {{1.0 / ex, -nxy / ex, -nxz / ex}, {-nxy / ex, 1.0 / ey, -nyz / ey},
                            {-nxz / ex, -nyz / ey, 1.0 / ez}};
    const double det = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1]) -
                       s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0]) +
                       s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
    const double minor = s[0][0] * s[1][1] - s[0][1] * s[1][0];
    if (!(minor > 0.0 && det > 1e-12 * s[0][0] * s[1][1] * s[2][2]))
        throw Error(
            "the Poisson ratios give no stable material (for an isotropic material, nu must be between -1 and 0.5)");
    std::array<double, 12> d{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            d[3 * i + j] = (s[j1][i1] * s[j2][i2] - s[j1][i2] * s[j2][i1]) / det;
        }
    d[9] = m[8];   // G_yz
    d[10] = m[7];  // G_xz
    d[11] = m[6];  // G_xy
    return d;
}

std::array<double, 576> hex_element(const std::array<double, 3>& size, const std::array<double, 12>& d) {
    std::array<double, 576> k;
    std::array<double, 192> load;
    hex_element_thermal(size, d, {0.0, 0.0, 0.0}, k, load);
    return k;
}

namespace {
// hex.rs:199-279: the element with its mode maps, mode_u (9 x 24) and mode_t (9 x 8): the modes'
// amplitudes are mode_t t - mode_u u.
void hex_element_modes(const std::array<double, 3>& size, const std::array<double, 12>& d,
                       const std::array<double, 3>& alpha, std::array<double, 576>& out, std::array<double, 192>& lout,
                       double x[9][24], double xt[9][8], double thermal[3]);
}  // namespace

void hex_element_thermal(const std::array<double, 3>& size, const std::array<double, 12>& d,
                         const std::array<double, 3>& alpha, std::array<double, 576>& out,
                         std::array<double, 192>& lout) {
    double x[9][24], xt[9][8], thermal[3];
    hex_element_modes(size, d, alpha, out, lout, x, xt, thermal);
}

void hex_strain_stress(const std::array<double, 3>& size, const std::array<double, 12>& d,
                       const std::array<double, 3>& alpha, const double xi[3], const double u[24], const double t[8],
                       double strain[6], double stress[6]) {
    std::array<double, 576> k;
    std::array<double, 192> load;
    double x[9][24], xt[9][8], thermal[3];
    hex_element_modes(size, d, alpha, k, load, x, xt, thermal);
    double all[11][3];
    for (int n = 0; n < 8; ++n)
        for (int c = 0; c < 3; ++c) all[n][c] = u[3 * n + c];
    for (int i = 0; i < 9; ++i) {
        double a = 0.0;
        for (int c = 0; c < 8; ++c) a += xt[i][c] * t[c];
        for (int c = 0; c < 24; ++c) a -= x[i][c] * u[c];
        all[8 + i / 3][i % 3] = a;
    }
    double g[11][3];
    gradients11(size, xi, g);
    for (int i = 0; i < 6; ++i) strain[i] = 0.0;
    for (int i = 0; i < 11; ++i) {  // hex.rs:180-193
        for (int dd = 0; dd < 3; ++dd) strain[dd] += g[i][dd] * all[i][dd];
        for (int sidx = 0; sidx < 3; ++sidx) {
            const int a = SHEAR_AXES[sidx][0], b = SHEAR_AXES[sidx][1];
            strain[3 + sidx] += g[i][a] * all[i][b] + g[i][b] * all[i][a];
        }
    }
    double dt = 0.0;  // the temperature change at xi: the corner functions over t
    for (int c = 0; c < 8; ++c) {
        double v = 1.0;
        for (int dd = 0; dd < 3; ++dd) v *= CORNER[c][dd] == 1 ? 0.5 * (1.0 + xi[dd]) : 0.5 * (1.0 - xi[dd]);
        dt += v * t[c];
    }
    const double* n = d.data();
    for (int i = 0; i < 3; ++i) stress[i] = n[3 * i] * strain[0] + n[3 * i + 1] * strain[1] + n[3 * i + 2] * strain[2];
    for (int sidx = 0; sidx < 3; ++sidx) stress[3 + sidx] = d[9 + sidx] * strain[3 + sidx];
    for (int i = 0; i < 3; ++i) stress[i] -= thermal[i] * dt;
}

namespace {
void hex_element_modes(const std::array<double, 3>& size, const std::array<double, 12>& d,
                       const std::array<double, 3>& alpha, std::array<double, 576>& out, std::array<double, 192>& lout,
                       double x[9][24], double xt[9][8], double thermal[3]) {
    const double* normal = d.data();
    const double* shear = d.data() + 9;
    // the thermal stress of a unit temperature change: D alpha on the normal stresses (a sum from 0)
    for (int i = 0; i < 3; ++i) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t += normal[3 * i + j] * alpha[j];
        thermal[i] = t;
    }
    std::vector<double> lbuf(33 * 8, 0.0);
    auto Lm = [&](int r, int c) -> double& { return lbuf[(size_t)r * 8 + c]; };
    std::vector<double> kbuf(33 * 33, 0.0);
    auto K = [&](int r, int c) -> double& { return kbuf[(size_t)r * 33 + c]; };
    const double jac = size[0] * size[1] * size[2] / 8.0;
    const double p = 1.0 / std::sqrt(3.0);
    for (int gp = 0; gp < 8; ++gp) {
        const double xi[3] = {CORNER[gp][0] == 1 ? p : -p, CORNER[gp][1] == 1 ? p : -p, CORNER[gp][2] == 1 ? p : -p};
        double g[11][3];
        gradients11(size, xi, g);
        double nshape[8];  // the corner functions at xi (hex.rs:148-163)
        for (int c = 0; c < 8; ++c) {
            double v = 1.0;
            for (int dd = 0; dd < 3; ++dd) v *= CORNER[c][dd] == 1 ? 0.5 * (1.0 + xi[dd]) : 0.5 * (1.0 - xi[dd]);
            nshape[c] = v;
        }
        for (int i = 0; i < 11; ++i) {
            const double* gi = g[i];
            for (int j = 0; j < 11; ++j) {
                const double* gj = g[j];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double v = gi[a] * gj[b] * normal[3 * a + b];
                        if (a == b) {
                            for (int s = 0; s < 3; ++s) {
                                const int sp = SHEAR_AXES[s][0], sq = SHEAR_AXES[s][1];
                                if (sp == a) v += gi[sq] * gj[sq] * shear[s];
                                else if (sq == a) v += gi[sp] * gj[sp] * shear[s];
                            }
                        } else {
                            v += gi[b] * gj[a] * shear[3 - a - b];
                        }
                        K(3 * i + a, 3 * j + b) += jac * v;
                    }
            }
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 8; ++c) Lm(3 * i + a, c) += jac * gi[a] * thermal[a] * nshape[c];
        }
    }
    // the modes condensed: X = K_mm^-1 K_mu, then K_uu - K_um X
    double kmm[9][9], l[9][9];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) kmm[i][j] = K(24 + i, 24 + j);
    factor9(kmm, l);
    for (int c = 0; c < 24; ++c) {
        double col[9];
        for (int i = 0; i < 9; ++i) col[i] = K(24 + i, c);
        solve9(l, col);
        for (int i = 0; i < 9; ++i) x[i][c] = col[i];
    }
    for (int c = 0; c < 8; ++c) {
        double col[9];
        for (int i = 0; i < 9; ++i) col[i] = Lm(24 + i, c);
        solve9(l, col);
        for (int i = 0; i < 9; ++i) xt[i][c] = col[i];
    }
    for (int r = 0; r < 24; ++r) {
        for (int c = 0; c < 24; ++c) {
            double v = K(r, c);
            for (int i = 0; i < 9; ++i) v -= K(r, 24 + i) * x[i][c];
            out[(size_t)r * 24 + c] = v;
        }
        for (int c = 0; c < 8; ++c) {
            double v = Lm(r, c);
            for (int i = 0; i < 9; ++i) v -= K(r, 24 + i) * xt[i][c];
            lout[(size_t)r * 8 + c] = v;
        }
    }
}
}  // namespace

std::vector<double> box_planes(const BoxModel& m, int d) {
    std::vector<double> p((size_t)m.in.n[d] + 1);
    for (size_t i = 0; i < p.size(); ++i) p[i] = m.h[d] * (double)i;
    return p;
}

HexBoxSystem hex_box_system(const BoxModel& m) {
    if (m.in.kind != Kind::ElasticHex8) throw Error("internal: a hex system of another kind");
    HexBoxSystem sys;
    const int nx = m.in.n[0], ny = m.in.n[1];
    const size_t cells = m.cells();
    std::array<std::vector<double>, 3> planes;
    for (int d = 0; d < 3; ++d) planes[d] = box_planes(m, d);
    // the element of each cell, computed once for each distinct (sizes, material)
    std::vector<uint32_t> elem_of(cells);
    std::vector<std::array<double, 576>> elems;
    std::map<std::array<uint64_t, 5>, uint32_t> seen;
    for (size_t e = 0; e < cells; ++e) {
        const size_t c[3] = {e % nx, (e / nx) % ny, e / ((size_t)nx * ny)};
        std::array<double, 3> size;
        for (int d = 0; d < 3; ++d) size[d] = planes[d][c[d] + 1] - planes[d][c[d]];
        const double ym = m.cell_k[e], nu = m.cell_nu[e];
        const std::array<uint64_t, 5> key = {bits_of(size[0]), bits_of(size[1]), bits_of(size[2]), bits_of(ym), bits_of(nu)};
        auto it = seen.find(key);
        if (it == seen.end()) {
            it = seen.emplace(key, (uint32_t)elems.size()).first;
            elems.push_back(hex_element(size, isotropic_stiffness(ym, nu)));
            const double hi = std::max({size[0], size[1], size[2]}), lo = std::min({size[0], size[1], size[2]});
            sys.aspect = std::max(sys.aspect, hi / lo);
        }
        elem_of[e] = it->second;
    }
    // the free index of each unknown of the reference numbering
    const size_t gnodes = m.grid_nodes();
    std::vector<uint32_t> point_of(gnodes);
    for (size_t p = 0; p < m.points; ++p) point_of[m.point_grid[p]] = (uint32_t)p;
    std::vector<int64_t> renum(m.unknowns(), -1);
    for (size_t f = 0; f < m.free.size(); ++f) renum[(size_t)m.free[f]] = (int64_t)f;
    const size_t nf = m.free.size();
    const int gx = m.gn[0], gy = m.gn[1], gz = m.gn[2];
    // Row of free unknown f: every free unknown of the 3 x 3 x 3 grid points around it, by column;
    // the value of an entry sums the cells that hold both points in increasing cell order, each
    // adding its element's entry, as add_element (sparse.rs:64-86) adds cell after cell.
    auto row = [&](size_t f, std::vector<std::pair<uint32_t, double>>& out) {
        out.clear();
        const int64_t u = m.free[f];
        const size_t p = (size_t)u / 3;
        const int c = (int)(u % 3);
        const uint32_t g = m.point_grid[p];
        const int q[3] = {(int)(g % gx), (int)((g / gx) % gy), (int)(g / ((size_t)gx * gy))};
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int o[3] = {q[0] + dx, q[1] + dy, q[2] + dz};
                    if (o[0] < 0 || o[0] >= gx || o[1] < 0 || o[1] >= gy || o[2] < 0 || o[2] >= gz) continue;
                    const size_t op = point_of[(size_t)o[0] + (size_t)gx * ((size_t)o[1] + (size_t)gy * o[2])];
                    // the cells holding both points: along each axis, the cells [max - 1, min] in range
                    int lo[3], hi[3];
                    for (int d = 0; d < 3; ++d) {
                        lo[d] = std::max(std::max(q[d], o[d]) - 1, 0);
                        hi[d] = std::min(std::min(q[d], o[d]), m.in.n[d] - 1);
                    }
                    for (int cc = 0; cc < 3; ++cc) {
                        const int64_t col = renum[3 * op + cc];
                        if (col < 0) continue;
                        double v = 0.0;
                        for (int k = lo[2]; k <= hi[2]; ++k)
                            for (int j = lo[1]; j <= hi[1]; ++j)
                                for (int i = lo[0]; i <= hi[0]; ++i) {
                                    const size_t e = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                                    const int a = CORNER_OF_BITS[(q[0] - i) + 2 * (q[1] - j) + 4 * (q[2] - k)];
                                    const int b = CORNER_OF_BITS[(o[0] - i) + 2 * (o[1] - j) + 4 * (o[2] - k)];
                                    v += elems[elem_of[e]][(size_t)(3 * a + c) * 24 + 3 * b + cc];
                                }
                        out.push_back({(uint32_t)col, v});
                    }
                }
        std::sort(out.begin(), out.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    };
    std::vector<uint64_t> count(nf + 1, 0);
    parallel_rows(nf, [&](size_t lo, size_t hi) {
        std::vector<std::pair<uint32_t, double>> r;
        for (size_t f = lo; f < hi; ++f) {
            row(f, r);
            count[f + 1] = r.size();
        }
    });
    sys.a.n = nf;
    sys.a.ptr.assign(nf + 1, 0);
    for (size_t f = 0; f < nf; ++f) sys.a.ptr[f + 1] = sys.a.ptr[f] + count[f + 1];
    sys.a.idx.resize(sys.a.ptr[nf]);
    sys.a.val.resize(sys.a.ptr[nf]);
    parallel_rows(nf, [&](size_t lo, size_t hi) {
        std::vector<std::pair<uint32_t, double>> r;
        for (size_t f = lo; f < hi; ++f) {
            row(f, r);
            for (size_t k = 0; k < r.size(); ++k) {
                sys.a.idx[sys.a.ptr[f] + k] = r[k].first;
                sys.a.val[sys.a.ptr[f] + k] = r[k].second;
            }
        }
    });
    // heat.rs:1219-1242 with the fixed values 0: the load of each free unknown
    sys.rhs.resize(nf);
    for (size_t f = 0; f < nf; ++f) sys.rhs[f] = m.load[(size_t)m.free_grid[f]];
    sys.level.planes = planes;
    sys.level.points.resize(nf);
    sys.level.comp.resize(nf);
    for (size_t f = 0; f < nf; ++f) {
        const uint32_t g = m.point_grid[(size_t)m.free[f] / 3];
        sys.level.points[f] = {g % (uint32_t)gx, (g / (uint32_t)gx) % (uint32_t)gy, g / ((uint32_t)gx * (uint32_t)gy)};
        sys.level.comp[f] = (uint8_t)(m.free[f] % 3);
    }
    sys.cells.origin.resize(cells);
    sys.cells.k.resize(cells);
    for (size_t e = 0; e < cells; ++e) {
        sys.cells.origin[e] = {(uint32_t)(e % nx), (uint32_t)((e / nx) % ny), (uint32_t)(e / ((size_t)nx * ny))};
        sys.cells.k[e] = elems[elem_of[e]];
    }
    return sys;
}

}  // namespace fem
