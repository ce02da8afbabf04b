// TEST INFRASTRUCTURE ONLY (the oracle). Restates the element routines of crates/fem/src:
// heat.rs (conduction, tet_conduction, tet_gradients), hex.rs (Orthotropic), tet10.rs (Tet10,
// Quadratic) and prism.rs (split_prism).
#include <algorithm>
#include <cmath>
#include <map>

#include "oracle.hpp"

namespace orc {

// heat.rs:1246-1278 (BITS in VTK hexahedron order, hex.rs:34-43)
static const usize HEX_BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

std::array<double, 64> conduction(Vec3 s, Vec3 k) {
    auto stiff = [](double h, usize a, usize b) { return a == b ? 1.0 / h : -1.0 / h; };
    auto mass = [](double h, usize a, usize b) { return a == b ? h / 3.0 : h / 6.0; };
    std::array<double, 64> ke{};
    for (usize p = 0; p < 8; ++p)
        for (usize q = 0; q < 8; ++q) {
            double v = 0.0;
            for (usize d = 0; d < 3; ++d) {
                double term = k[d];
                for (usize e = 0; e < 3; ++e)
                    term *= e == d ? stiff(s[e], HEX_BITS[p][e], HEX_BITS[q][e])
                                   : mass(s[e], HEX_BITS[p][e], HEX_BITS[q][e]);
                v += term;
            }
            ke[p * 8 + q] = v;
        }
    return ke;
}

// heat.rs:972-980
std::array<double, 16> tet_conduction(const std::array<Vec3, 4>& g, double v, Vec3 k) {
    std::array<double, 16> ke{};
    for (usize i = 0; i < 4; ++i)
        for (usize j = 0; j < 4; ++j) {
            double s = 0.0;
            for (usize d = 0; d < 3; ++d) s += k[d] * g[i][d] * g[j][d];
            ke[i * 4 + j] = v * s;
        }
    return ke;
}

// heat.rs:983-1000
std::pair<std::array<Vec3, 4>, double> tet_gradients(const std::array<Vec3, 4>& p) {
    Vec3 e[3];
    for (usize k = 0; k < 3; ++k)
        for (usize d = 0; d < 3; ++d) e[k][d] = p[k + 1][d] - p[0][d];
    auto cross = [](const Vec3& a, const Vec3& b) -> Vec3 {
        return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    };
    const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                       e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                       e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    const Vec3 rows[3] = {cross(e[1], e[2]), cross(e[2], e[0]), cross(e[0], e[1])};
    std::array<Vec3, 4> g{};
    for (usize r = 0; r < 3; ++r)
        for (usize d = 0; d < 3; ++d) g[r + 1][d] = rows[r][d] / det;
    for (usize d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
    return {g, det / 6.0};
}

// hex.rs:58-64
Orthotropic Orthotropic::isotropic(double e, double nu) {
    const double g = e / (2.0 * (1.0 + nu));
    return {{e, e, e}, {nu, nu, nu}, {g, g, g}};
}

// hex.rs:69-115
Stiffness Orthotropic::stiffness() const {
    const double ex = e[0], ey = e[1], ez = e[2];
    const double nxy = nu[0], nxz = nu[1], nyz = nu[2];
    bool ok = true;
    for (double v : e) ok = ok && std::isfinite(v) && v > 0.0;
    for (double v : g) ok = ok && std::isfinite(v) && v > 0.0;
    bool nu_ok = true;
    for (double v : nu) nu_ok = nu_ok && std::isfinite(v);
    if (!ok || !nu_ok) throw Error("the moduli must be above 0 and the Poisson ratios finite");
    const double s[3][3] = {{1.0 / ex, -nxy / ex, -nxz / ex}, {-nxy / ex, 1.0 / ey, -nyz / ey}, {-nxz / ex, -nyz / ey, 1.0 / ez}};
    const double det = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1]) -
                       s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0]) +
                       s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
    const double minor = s[0][0] * s[1][1] - s[0][1] * s[1][0];
    if (!(minor > 0.0 && det > 1e-12 * s[0][0] * s[1][1] * s[2][2]))
        throw Error(
            "the Poisson ratios give no stable material (for an isotropic material, nu must be between -1 and 0.5)");
    Stiffness out{};
    for (usize i = 0; i < 3; ++i)
        for (usize j = 0; j < 3; ++j) {
            const usize i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            const usize j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            out.normal[i][j] = (s[j1][i1] * s[j2][i2] - s[j1][i2] * s[j2][i1]) / det;
        }
    out.shear[0] = g[2];  // G_yz
    out.shear[1] = g[1];  // G_xz
    out.shear[2] = g[0];  // G_xy
    return out;
}

// tet10.rs:38 and 41
const usize TET10_EDGES[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
static const usize SHEAR[3][2] = {{1, 2}, {0, 2}, {0, 1}};

// tet10.rs:44-48
std::array<std::array<double, 4>, 4> tet10_gauss() {
    const double a = 0.5854101966249685;
    const double b = 0.1381966011250105;
    std::array<std::array<double, 4>, 4> out{};
    for (usize q = 0; q < 4; ++q)
        for (usize k = 0; k < 4; ++k) out[q][k] = k == q ? a : b;
    return out;
}

// tet10.rs:71-74
Tet10 Tet10::make(const std::array<Vec3, 4>& p) {
    auto [g, v] = tet_gradients(p);
    Tet10 t;
    t.g = g;
    t.volume = v;
    return t;
}

// tet10.rs:77-87
std::array<Vec3, 10> Tet10::gradients(const std::array<double, 4>& l) const {
    std::array<Vec3, 10> out{};
    for (usize i = 0; i < 4; ++i)
        for (usize d = 0; d < 3; ++d) out[i][d] = (4.0 * l[i] - 1.0) * g[i][d];
    for (usize k = 0; k < 6; ++k) {
        const usize i = TET10_EDGES[k][0], j = TET10_EDGES[k][1];
        for (usize d = 0; d < 3; ++d) out[4 + k][d] = 4.0 * (l[i] * g[j][d] + l[j] * g[i][d]);
    }
    return out;
}

// The B^T D B term of unknowns (i, a) and (j, b) for the gradients gi, gj (tet10.rs:99-112).
static double bdb(const Vec3& gi, const Vec3& gj, usize a, usize b, const Stiffness& d) {
    double v = gi[a] * gj[b] * d.normal[a][b];
    if (a == b) {
        for (usize s = 0; s < 3; ++s) {
            const usize p = SHEAR[s][0], q = SHEAR[s][1];
            if (p == a) v += gi[q] * gj[q] * d.shear[s];
            else if (q == a) v += gi[p] * gj[p] * d.shear[s];
        }
    } else {
        v += gi[b] * gj[a] * d.shear[3 - a - b];
    }
    return v;
}

// tet10.rs:90-120
std::vector<double> Tet10::stiffness(const Stiffness& d) const {
    std::vector<double> k(900, 0.0);
    const double w = volume / 4.0;
    for (const auto& l : tet10_gauss()) {
        const auto gr = gradients(l);
        for (usize i = 0; i < 10; ++i)
            for (usize j = 0; j < 10; ++j)
                for (usize a = 0; a < 3; ++a)
                    for (usize b = 0; b < 3; ++b) k[(3 * i + a) * 30 + 3 * j + b] += w * bdb(gr[i], gr[j], a, b, d);
    }
    return k;
}

// tet10.rs:170-212
std::vector<double> Tet10::mass(double density) const {
    auto contains = [](const usize* e, usize c) { return e[0] == c || e[1] == c; };
    auto scalar = [&](usize i, usize j) -> double {
        const bool ci = i < 4, cj = j < 4;
        if (ci && cj) return i == j ? 6.0 : 1.0;
        if (ci != cj) {
            const usize c = ci ? i : j;
            const usize e = ci ? j - 4 : i - 4;
            return contains(TET10_EDGES[e], c) ? -4.0 : -6.0;
        }
        const usize* a = TET10_EDGES[i - 4];
        const usize* b = TET10_EDGES[j - 4];
        int shared = (contains(b, a[0]) ? 1 : 0) + (contains(b, a[1]) ? 1 : 0);
        return shared == 2 ? 32.0 : shared == 1 ? 16.0 : 8.0;
    };
    const double w = density * volume / 420.0;
    std::vector<double> m(900, 0.0);
    for (usize i = 0; i < 10; ++i)
        for (usize j = 0; j < 10; ++j) {
            const double v = w * scalar(i, j);
            for (usize a = 0; a < 3; ++a) m[(3 * i + a) * 30 + 3 * j + a] = v;
        }
    return m;
}

std::vector<double> tet4_stiffness(const std::array<Vec3, 4>& g, double volume, const Stiffness& d) {
    std::vector<double> k(144, 0.0);
    for (usize i = 0; i < 4; ++i)
        for (usize j = 0; j < 4; ++j)
            for (usize a = 0; a < 3; ++a)
                for (usize b = 0; b < 3; ++b) k[(3 * i + a) * 12 + 3 * j + b] = volume * bdb(g[i], g[j], a, b, d);
    return k;
}

// prism.rs:131-154
static const usize ROTATE[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 2, 0, 4, 5, 3}, {2, 0, 1, 5, 3, 4},
                                   {3, 5, 4, 0, 2, 1}, {4, 3, 5, 1, 0, 2}, {5, 4, 3, 2, 1, 0}};

std::array<std::array<usize, 4>, 3> split_prism(const std::array<std::uint64_t, 6>& g) {
    usize m = 0;  // min_by_key returns the first minimum
    for (usize k = 1; k < 6; ++k)
        if (g[k] < g[m]) m = k;
    const usize* r = ROTATE[m];
    auto gr = [&](usize k) { return g[r[k]]; };
    static const usize A[3][4] = {{0, 1, 2, 5}, {0, 1, 5, 4}, {0, 4, 5, 3}};
    static const usize B[3][4] = {{0, 1, 2, 4}, {0, 4, 2, 5}, {0, 4, 5, 3}};
    const auto& t = std::min(gr(1), gr(5)) < std::min(gr(2), gr(4)) ? A : B;
    std::array<std::array<usize, 4>, 3> out{};
    for (usize i = 0; i < 3; ++i)
        for (usize j = 0; j < 4; ++j) out[i][j] = r[t[i][j]];
    return out;
}

// tet10.rs:244-272
Quadratic Quadratic::make(const std::vector<Vec3>& points, const std::vector<std::array<u32, 4>>& tets) {
    Quadratic q;
    q.corners = points.size();
    std::map<std::pair<u32, u32>, u32> of;
    q.tets.reserve(tets.size());
    for (const auto& tet : tets) {
        std::array<u32, 10> t{};
        for (usize c = 0; c < 4; ++c) t[c] = tet[c];
        for (usize k = 0; k < 6; ++k) {
            const u32 x = tet[TET10_EDGES[k][0]], y = tet[TET10_EDGES[k][1]];
            const u32 a = std::min(x, y), b = std::max(x, y);
            auto it = of.find({a, b});
            if (it == of.end()) {
                q.edges.push_back({a, b});
                it = of.emplace(std::make_pair(a, b), (u32)(q.corners + q.edges.size() - 1)).first;
            }
            t[4 + k] = it->second;
        }
        q.tets.push_back(t);
    }
    q.points = points;
    for (const auto& e : q.edges) {
        const Vec3& p = points[e[0]];
        const Vec3& r = points[e[1]];
        q.points.push_back({0.5 * (p[0] + r[0]), 0.5 * (p[1] + r[1]), 0.5 * (p[2] + r[2])});
    }
    return q;
}


// ------------------------------------------------------------------ hex.rs:146-260 (Hex::new, k)
namespace {
// hex.rs:146-161: the gradients of the 8 corner functions and the 3 modes at xi.
std::array<Vec3, 11> hex_gradients(const Vec3& size, const Vec3& xi) {
    std::array<Vec3, 11> g{};
    for (usize n = 0; n < 8; ++n) {
        double s[3], f[3];
        for (usize d = 0; d < 3; ++d) {
            s[d] = HEX_BITS[n][d] == 1 ? 1.0 : -1.0;
            f[d] = 0.5 * (1.0 + s[d] * xi[d]);
        }
        for (usize d = 0; d < 3; ++d) {
            const usize d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            g[n][d] = s[d] / size[d] * f[d1] * f[d2];
        }
    }
    for (usize d = 0; d < 3; ++d) g[8 + d][d] = -4.0 * xi[d] / size[d];
    return g;
}

// hex.rs:364-382
std::array<std::array<double, 9>, 9> cholesky9(const std::array<std::array<double, 9>, 9>& a) {
    std::array<std::array<double, 9>, 9> l{};
    for (usize j = 0; j < 9; ++j) {
        double d = a[j][j];
        for (usize k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
        d = std::sqrt(d);
        l[j][j] = d;
        for (usize i = j + 1; i < 9; ++i) {
            double v = a[i][j];
            for (usize k = 0; k < j; ++k) v -= l[i][k] * l[j][k];
            l[i][j] = v / d;
        }
    }
    return l;
}

// hex.rs:384-398
std::array<double, 9> solve9(const std::array<std::array<double, 9>, 9>& l, std::array<double, 9> b) {
    for (usize i = 0; i < 9; ++i) {
        for (usize k = 0; k < i; ++k) b[i] -= l[i][k] * b[k];
        b[i] /= l[i][i];
    }
    for (usize i = 9; i-- > 0;) {
        for (usize k = i + 1; k < 9; ++k) b[i] -= l[k][i] * b[k];
        b[i] /= l[i][i];
    }
    return b;
}
}  // namespace

std::array<double, 8> hex_shape(const Vec3& xi) {  // hex.rs:164-178
    std::array<double, 8> out{};
    for (usize n = 0; n < 8; ++n) {
        double p = 1.0;
        for (usize d = 0; d < 3; ++d) p *= HEX_BITS[n][d] == 1 ? 0.5 * (1.0 + xi[d]) : 0.5 * (1.0 - xi[d]);
        out[n] = p;
    }
    return out;
}

std::array<double, 576> hex_stiffness(const Vec3& size, const Stiffness& dm) {
    static const usize SHEAR_AX[3][2] = {{1, 2}, {0, 2}, {0, 1}};  // hex.rs:141
    std::array<std::array<double, 33>, 33> k{};
    const double jac = size[0] * size[1] * size[2] / 8.0;
    const double p = 1.0 / std::sqrt(3.0);
    for (usize q = 0; q < 8; ++q) {  // hex.rs:193-196: the 2-point Gauss points in BITS order
        const Vec3 xi{HEX_BITS[q][0] == 1 ? p : -p, HEX_BITS[q][1] == 1 ? p : -p, HEX_BITS[q][2] == 1 ? p : -p};
        const auto g = hex_gradients(size, xi);
        for (usize i = 0; i < 11; ++i)
            for (usize j = 0; j < 11; ++j)
                for (usize a = 0; a < 3; ++a)
                    for (usize b = 0; b < 3; ++b) {
                        double v = g[i][a] * g[j][b] * dm.normal[a][b];
                        if (a == b) {
                            for (usize s = 0; s < 3; ++s) {
                                if (SHEAR_AX[s][0] == a) v += g[i][SHEAR_AX[s][1]] * g[j][SHEAR_AX[s][1]] * dm.shear[s];
                                else if (SHEAR_AX[s][1] == a) v += g[i][SHEAR_AX[s][0]] * g[j][SHEAR_AX[s][0]] * dm.shear[s];
                            }
                        } else {
                            const usize s = 3 - a - b;
                            v += g[i][b] * g[j][a] * dm.shear[s];
                        }
                        k[3 * i + a][3 * j + b] += jac * v;
                    }
    }
    // condense the 9 modes (hex.rs:238-279)
    std::array<std::array<double, 9>, 9> kmm{};
    for (usize i = 0; i < 9; ++i)
        for (usize j = 0; j < 9; ++j) kmm[i][j] = k[24 + i][24 + j];
    const auto chol = cholesky9(kmm);
    std::array<double, 216> mode_u{};
    for (usize c = 0; c < 24; ++c) {
        std::array<double, 9> rhs{};
        for (usize i = 0; i < 9; ++i) rhs[i] = k[24 + i][c];
        const auto col = solve9(chol, rhs);
        for (usize i = 0; i < 9; ++i) mode_u[i * 24 + c] = col[i];
    }
    std::array<double, 576> kc{};
    for (usize r = 0; r < 24; ++r)
        for (usize c = 0; c < 24; ++c) {
            double v = k[r][c];
            for (usize i = 0; i < 9; ++i) v -= k[r][24 + i] * mode_u[i * 24 + c];
            kc[r * 24 + c] = v;
        }
    return kc;
}

}  // namespace orc
