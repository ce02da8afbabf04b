// TEST INFRASTRUCTURE ONLY (the oracle). A C surface over the restatement, so the Python tests
// and the CPU-baseline leg of the benchmark can drive it with ctypes. Handles own their data;
// each `*_free` releases one. Errors come back as the reference's text in `err`.
#include <cstring>
#include <memory>

#include "oracle.hpp"

using namespace orc;

namespace {
void set_err(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}
std::vector<Vec3> xyz_of(const double* p, size_t n) {
    std::vector<Vec3> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return v;
}
}  // namespace

extern "C" {

// ------------------------------------------------------------------ Csr
void* orc_csr_pattern(size_t n, size_t ne, size_t npe, const uint32_t* conn) {
    std::vector<usize> eptr(ne + 1);
    for (size_t e = 0; e <= ne; ++e) eptr[e] = e * npe;
    std::vector<u32> c(conn, conn + ne * npe);
    return new Csr(Csr::pattern(n, eptr, c));
}
void* orc_csr_from(size_t n, const int64_t* ptr, const uint32_t* idx, const double* val) {
    auto* a = new Csr;
    a->n = n;
    a->ptr.assign(ptr, ptr + n + 1);
    a->idx.assign(idx, idx + ptr[n]);
    a->val.assign(val, val + ptr[n]);
    return a;
}
int orc_csr_add_element(void* h, const uint32_t* nodes, size_t m, const double* ke, char* err, size_t errlen) {
    try {
        static_cast<Csr*>(h)->add_element(nodes, m, ke);
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}
size_t orc_csr_n(void* h) { return static_cast<Csr*>(h)->n; }
size_t orc_csr_nnz(void* h) { return static_cast<Csr*>(h)->idx.size(); }
void orc_csr_get(void* h, int64_t* ptr, uint32_t* idx, double* val) {
    const Csr* a = static_cast<Csr*>(h);
    for (size_t i = 0; i <= a->n; ++i) ptr[i] = (int64_t)a->ptr[i];
    std::memcpy(idx, a->idx.data(), a->idx.size() * sizeof(uint32_t));
    std::memcpy(val, a->val.data(), a->val.size() * sizeof(double));
}
void orc_csr_mul(void* h, const double* x, double* y) { static_cast<Csr*>(h)->mul(x, y); }
void orc_csr_diagonal(void* h, double* out) {
    const auto d = static_cast<Csr*>(h)->diagonal();
    std::memcpy(out, d.data(), d.size() * sizeof(double));
}
void* orc_csr_submatrix(void* h, size_t nkeep, const int64_t* keep) {
    std::vector<usize> k(keep, keep + nkeep);
    return new Csr(static_cast<Csr*>(h)->submatrix(k));
}
void orc_csr_free(void* h) { delete static_cast<Csr*>(h); }

// ------------------------------------------------------------------ Rect
void* orc_rect_from(size_t rows, size_t cols, const int64_t* ptr, const uint32_t* idx, const double* val) {
    auto* p = new Rect;
    p->rows = rows;
    p->cols = cols;
    p->ptr.assign(ptr, ptr + rows + 1);
    p->idx.assign(idx, idx + ptr[rows]);
    p->val.assign(val, val + ptr[rows]);
    return p;
}
size_t orc_rect_rows(void* h) { return static_cast<Rect*>(h)->rows; }
size_t orc_rect_cols(void* h) { return static_cast<Rect*>(h)->cols; }
size_t orc_rect_nnz(void* h) { return static_cast<Rect*>(h)->idx.size(); }
void orc_rect_get(void* h, int64_t* ptr, uint32_t* idx, double* val) {
    const Rect* p = static_cast<Rect*>(h);
    for (size_t i = 0; i <= p->rows; ++i) ptr[i] = (int64_t)p->ptr[i];
    std::memcpy(idx, p->idx.data(), p->idx.size() * sizeof(uint32_t));
    std::memcpy(val, p->val.data(), p->val.size() * sizeof(double));
}
void orc_rect_mul(void* h, const double* x, double* y) { static_cast<Rect*>(h)->mul(x, y); }
void orc_rect_mul_transpose(void* h, const double* x, double* y) { static_cast<Rect*>(h)->mul_transpose(x, y); }
void* orc_rect_galerkin(void* h, void* a) { return new Csr(static_cast<Rect*>(h)->galerkin(*static_cast<Csr*>(a))); }
void orc_rect_free(void* h) { delete static_cast<Rect*>(h); }

// ------------------------------------------------------------------ ordering, factor
void orc_nested_dissection(void* a, const double* xyz, int64_t* order) {
    const Csr* m = static_cast<Csr*>(a);
    const auto o = nested_dissection(*m, xyz_of(xyz, m->n));
    for (size_t k = 0; k < o.size(); ++k) order[k] = (int64_t)o[k];
}
void* orc_symbolic(void* a, const int64_t* order) {
    const Csr* m = static_cast<Csr*>(a);
    std::vector<usize> o(order, order + m->n);
    return new Symbolic(Symbolic::analyse(*m, o));
}
size_t orc_symbolic_supernodes(void* h) { return static_cast<Symbolic*>(h)->supernodes(); }
uint64_t orc_symbolic_entries(void* h) { return static_cast<Symbolic*>(h)->entries(); }
double orc_symbolic_flops(void* h) { return static_cast<Symbolic*>(h)->flops(); }
uint64_t orc_symbolic_largest_front(void* h) { return static_cast<Symbolic*>(h)->largest_front(); }
void orc_symbolic_perm(void* h, int64_t* out) {
    const Symbolic* s = static_cast<Symbolic*>(h);
    for (size_t k = 0; k < s->n; ++k) out[k] = (int64_t)s->perm[k];
}
size_t orc_symbolic_rows_len(void* h) { return static_cast<Symbolic*>(h)->rows.size(); }
void orc_symbolic_structure(void* h, int64_t* start, int64_t* rows_ptr, uint32_t* rows, int64_t* parent) {
    const Symbolic* s = static_cast<Symbolic*>(h);
    for (size_t k = 0; k < s->start.size(); ++k) start[k] = (int64_t)s->start[k];
    for (size_t k = 0; k < s->rows_ptr.size(); ++k) rows_ptr[k] = (int64_t)s->rows_ptr[k];
    for (size_t k = 0; k < s->rows.size(); ++k) rows[k] = s->rows[k];
    for (size_t k = 0; k < s->parent.size(); ++k) parent[k] = s->parent[k] == (usize)-1 ? -1 : (int64_t)s->parent[k];
}
void orc_symbolic_free(void* h) { delete static_cast<Symbolic*>(h); }
void* orc_cholesky(void* sym, void* a, char* err, size_t errlen) {
    try {
        return new Cholesky(Cholesky::factor(*static_cast<Symbolic*>(sym), *static_cast<Csr*>(a)));
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}
void orc_cholesky_solve(void* h, double* b) { static_cast<Cholesky*>(h)->solve(b); }
size_t orc_cholesky_entries(void* h) { return static_cast<Cholesky*>(h)->entries(); }
void orc_cholesky_free(void* h) { delete static_cast<Cholesky*>(h); }

// ------------------------------------------------------------------ elements
void orc_conduction(const double* s, const double* k, double* out) {
    const auto ke = conduction({s[0], s[1], s[2]}, {k[0], k[1], k[2]});
    std::memcpy(out, ke.data(), 64 * sizeof(double));
}
void orc_tet_gradients(const double* p, double* g, double* v) {
    const auto [gg, vv] = tet_gradients({Vec3{p[0], p[1], p[2]}, Vec3{p[3], p[4], p[5]}, Vec3{p[6], p[7], p[8]},
                                         Vec3{p[9], p[10], p[11]}});
    for (int i = 0; i < 4; ++i)
        for (int d = 0; d < 3; ++d) g[3 * i + d] = gg[i][d];
    *v = vv;
}
void orc_tet_conduction(const double* g, double v, const double* k, double* out) {
    std::array<Vec3, 4> gg;
    for (int i = 0; i < 4; ++i) gg[i] = {g[3 * i], g[3 * i + 1], g[3 * i + 2]};
    const auto ke = tet_conduction(gg, v, {k[0], k[1], k[2]});
    std::memcpy(out, ke.data(), 16 * sizeof(double));
}
int orc_orthotropic_stiffness(const double* e, const double* nu, const double* g, double* normal, double* shear,
                              char* err, size_t errlen) {
    try {
        Orthotropic o{{e[0], e[1], e[2]}, {nu[0], nu[1], nu[2]}, {g[0], g[1], g[2]}};
        const Stiffness s = o.stiffness();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) normal[3 * i + j] = s.normal[i][j];
        for (int i = 0; i < 3; ++i) shear[i] = s.shear[i];
        return 0;
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return 1;
    }
}
void orc_isotropic(double e, double nu, double* oe, double* onu, double* og) {
    const Orthotropic o = Orthotropic::isotropic(e, nu);
    for (int i = 0; i < 3; ++i) {
        oe[i] = o.e[i];
        onu[i] = o.nu[i];
        og[i] = o.g[i];
    }
}
static Stiffness stiff_of(const double* normal, const double* shear) {
    Stiffness s{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s.normal[i][j] = normal[3 * i + j];
    for (int i = 0; i < 3; ++i) s.shear[i] = shear[i];
    return s;
}
static std::array<Vec3, 4> four(const double* p) {
    return {Vec3{p[0], p[1], p[2]}, Vec3{p[3], p[4], p[5]}, Vec3{p[6], p[7], p[8]}, Vec3{p[9], p[10], p[11]}};
}
void orc_tet10_stiffness(const double* p, const double* normal, const double* shear, double* out) {
    const auto k = Tet10::make(four(p)).stiffness(stiff_of(normal, shear));
    std::memcpy(out, k.data(), 900 * sizeof(double));
}
void orc_tet10_mass(const double* p, double density, double* out) {
    const auto m = Tet10::make(four(p)).mass(density);
    std::memcpy(out, m.data(), 900 * sizeof(double));
}
void orc_tet4_stiffness(const double* p, const double* normal, const double* shear, double* out) {
    const auto [g, v] = tet_gradients(four(p));
    const auto k = tet4_stiffness(g, v, stiff_of(normal, shear));
    std::memcpy(out, k.data(), 144 * sizeof(double));
}
void orc_split_prism(const uint64_t* g, int64_t* out) {
    const auto t = split_prism({g[0], g[1], g[2], g[3], g[4], g[5]});
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = (int64_t)t[i][j];
}

// Quadratic mesh: returns the number of edges; `edges_out` (2 per edge) and `tets_out` (10 per
// tet) must hold enough room (6 edges per tet at most).
size_t orc_quadratic(size_t npts, const double* pts, size_t ntets, const uint32_t* tets, uint32_t* edges_out,
                     uint32_t* tets_out) {
    std::vector<std::array<u32, 4>> t(ntets);
    for (size_t e = 0; e < ntets; ++e) t[e] = {tets[4 * e], tets[4 * e + 1], tets[4 * e + 2], tets[4 * e + 3]};
    const Quadratic q = Quadratic::make(xyz_of(pts, npts), t);
    for (size_t k = 0; k < q.edges.size(); ++k) {
        edges_out[2 * k] = q.edges[k][0];
        edges_out[2 * k + 1] = q.edges[k][1];
    }
    for (size_t e = 0; e < ntets; ++e)
        for (int c = 0; c < 10; ++c) tets_out[10 * e + c] = q.tets[e][c];
    return q.edges.size();
}

void* orc_prolongation(size_t npts, const double* pts, size_t corners, size_t nedges, const uint32_t* edges,
                       size_t nfree, const int64_t* free, double* coarse_xyz) {
    std::vector<std::array<u32, 2>> ed(nedges);
    for (size_t k = 0; k < nedges; ++k) ed[k] = {edges[2 * k], edges[2 * k + 1]};
    std::vector<usize> f(free, free + nfree);
    auto [p, xyz] = prolongation(xyz_of(pts, npts), corners, ed, f);
    if (coarse_xyz)
        for (size_t i = 0; i < xyz.size(); ++i)
            for (int d = 0; d < 3; ++d) coarse_xyz[3 * i + d] = xyz[i][d];
    return new Rect(std::move(p));
}

// ------------------------------------------------------------------ structured box problems
static BoxSpec box_of(const int* n, const double* len) { return {n[0], n[1], n[2], len[0], len[1], len[2]}; }

void* orc_problem_heat(int kind_tet4, const int* n, const double* len, const double* k, const int* fixed,
                       const double* value, double source, char* err, size_t errlen) {
    try {
        HeatBoxSpec s;
        s.box = box_of(n, len);
        const size_t ncell = (size_t)n[0] * n[1] * n[2];
        s.k.assign(k, k + (kind_tet4 ? 6 * ncell : ncell));
        for (int f = 0; f < 6; ++f) {
            s.fixed[f] = fixed[f] != 0;
            s.value[f] = value[f];
        }
        s.source = source;
        return new Problem(kind_tet4 ? heat_tet4_box(s) : heat_hex8_box(s));
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}
void* orc_problem_elastic(int order, const int* n, const double* len, const double* e, const double* nu,
                          const int* clamped, const double* body, const double* traction, char* err, size_t errlen) {
    try {
        ElasticBoxSpec s;
        s.box = box_of(n, len);
        s.order = order;
        const size_t nt = 6 * (size_t)n[0] * n[1] * n[2];
        s.e.assign(e, e + nt);
        s.nu.assign(nu, nu + nt);
        for (int f = 0; f < 6; ++f) {
            s.clamped[f] = clamped[f] != 0;
            s.traction[f] = {traction[3 * f], traction[3 * f + 1], traction[3 * f + 2]};
        }
        s.body = {body[0], body[1], body[2]};
        return new Problem(elastic_tet_box(s));
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return nullptr;
    }
}
// Box-grid hex elasticity: E and nu per cell (elastic.rs:287-563).
void* orc_problem_elastic_hex(const int* n, const double* len, const double* e, const double* nu, const int* clamped,
                              const double* body, char* err, size_t errlen) {
    try {
        ElasticHexSpec s;
        s.box = box_of(n, len);
        const size_t nc = (size_t)n[0] * n[1] * n[2];
        s.e.assign(e, e + nc);
        s.nu.assign(nu, nu + nc);
        for (int f = 0; f < 6; ++f) s.clamped[f] = clamped[f] != 0;
        s.body = {body[0], body[1], body[2]};
        return new Problem(elastic_hex8_box(s));
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return nullptr;
    }
}
// The cells of a hex elasticity problem: each cell's lowest grid point (3) and matrix (576).
size_t orc_problem_cell_count(void* h) { return static_cast<Problem*>(h)->cell_matrices.origin.size(); }
void orc_problem_cell_matrices(void* h, uint32_t* origin, double* k) {
    const Problem* p = static_cast<Problem*>(h);
    for (size_t c = 0; c < p->cell_matrices.origin.size(); ++c) {
        for (int d = 0; d < 3; ++d) origin[3 * c + d] = p->cell_matrices.origin[c][d];
        std::copy(p->cell_matrices.k[c].begin(), p->cell_matrices.k[c].end(), k + 576 * c);
    }
}
// hex.rs:199-279: the condensed 24 x 24 element of a box cell of an isotropic material.
void orc_hex_stiffness(const double* size, double e, double nu, double* out, char* err, size_t errlen) {
    try {
        const auto k = hex_stiffness({size[0], size[1], size[2]}, Orthotropic::isotropic(e, nu).stiffness());
        std::copy(k.begin(), k.end(), out);
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
    }
}
// One coarsen and by_cells step (multigrid.rs:96-125) on a hex elasticity problem's reduced
// matrix, its grid level and its cells, with `all_free` the whole box free (no support): the coarse
// matrix (a Csr handle) and the coarse points and components (3 and 1 per coarse unknown).
void* orc_by_cells_step(void* h, uint32_t* coarse_points, uint8_t* coarse_comp, size_t* coarse_n) {
    const Problem* p = static_cast<Problem*>(h);
    const Csr a = p->k.submatrix(p->free);
    auto c = coarsen(p->unknowns.level, a);
    if (!c) return nullptr;
    ByCells bc = by_cells(p->unknowns.level, p->cell_matrices, c->coarse, c->kept);
    for (size_t i = 0; i < c->coarse.points.size(); ++i) {
        for (int d = 0; d < 3; ++d) coarse_points[3 * i + d] = c->coarse.points[i][d];
        coarse_comp[i] = c->coarse.comp[i];
    }
    *coarse_n = c->coarse.points.size();
    return new Csr(std::move(bc.a));
}
// Solver::new on a problem with its cells given (multigrid.rs:84-139 by_cells levels).
void* orc_solver_new_cells(void* h, int method, size_t crossover_unknowns, char* err, size_t errlen) {
    try {
        const Problem* p = static_cast<Problem*>(h);
        std::optional<Method> m;
        if (method == 0) m = Method::Direct;
        if (method == 1) m = Method::Multigrid;
        Crossover c;
        c.unknowns = crossover_unknowns;
        return new Solver(Solver::make(p->k.submatrix(p->free), p->unknowns, m, c, &p->cell_matrices));
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return nullptr;
    }
}
// sizes: n_points, comps, n_unknowns, n_free, npe, n_cells, corners, n_edges
void orc_problem_sizes(void* h, size_t* out) {
    const Problem* p = static_cast<Problem*>(h);
    out[0] = p->n_points;
    out[1] = p->comps;
    out[2] = p->n_points * p->comps;
    out[3] = p->free.size();
    out[4] = p->npe;
    out[5] = p->npe ? p->cells.size() / p->npe : 0;
    out[6] = p->corners;
    out[7] = p->edges.size();
}
void orc_problem_points(void* h, double* out) {
    const Problem* p = static_cast<Problem*>(h);
    for (size_t i = 0; i < p->points.size(); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = p->points[i][d];
}
void orc_problem_ijk(void* h, uint32_t* out) {
    const Problem* p = static_cast<Problem*>(h);
    for (size_t i = 0; i < p->ijk.size(); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = p->ijk[i][d];
}
void orc_problem_cells(void* h, uint32_t* out) {
    const Problem* p = static_cast<Problem*>(h);
    std::memcpy(out, p->cells.data(), p->cells.size() * sizeof(uint32_t));
}
void orc_problem_edges(void* h, uint32_t* out) {
    const Problem* p = static_cast<Problem*>(h);
    for (size_t k = 0; k < p->edges.size(); ++k) {
        out[2 * k] = p->edges[k][0];
        out[2 * k + 1] = p->edges[k][1];
    }
}
void orc_problem_free(void* h, int64_t* out) {
    const Problem* p = static_cast<Problem*>(h);
    for (size_t k = 0; k < p->free.size(); ++k) out[k] = (int64_t)p->free[k];
}
void orc_problem_rhs(void* h, double* out) {
    const Problem* p = static_cast<Problem*>(h);
    std::memcpy(out, p->rhs.data(), p->rhs.size() * sizeof(double));
}
void orc_problem_load(void* h, double* out) {
    const Problem* p = static_cast<Problem*>(h);
    std::memcpy(out, p->load.data(), p->load.size() * sizeof(double));
}
void orc_problem_fixed(void* h, uint8_t* is_fixed, double* value) {
    const Problem* p = static_cast<Problem*>(h);
    std::memcpy(is_fixed, p->is_fixed.data(), p->is_fixed.size());
    std::memcpy(value, p->fixed_value.data(), p->fixed_value.size() * sizeof(double));
}
void* orc_problem_matrix(void* h) { return new Csr(static_cast<Problem*>(h)->k); }
void* orc_problem_reduced(void* h) {
    const Problem* p = static_cast<Problem*>(h);
    return new Csr(p->k.submatrix(p->free));
}
void* orc_problem_prolongation(void* h) {
    const Problem* p = static_cast<Problem*>(h);
    if (p->unknowns.kind != Unknowns::Quadratic) return nullptr;
    return new Rect(p->unknowns.p);
}
size_t orc_problem_coarse_xyz(void* h, double* out) {
    const Problem* p = static_cast<Problem*>(h);
    if (out)
        for (size_t i = 0; i < p->unknowns.coarse.size(); ++i)
            for (int d = 0; d < 3; ++d) out[3 * i + d] = p->unknowns.coarse[i][d];
    return p->unknowns.coarse.size();
}
void orc_problem_free_handle(void* h) { delete static_cast<Problem*>(h); }

// ------------------------------------------------------------------ Solver
// method: -1 none (the crossover chooses), 0 Direct, 1 Multigrid.
void* orc_solver_new(void* h, int method, size_t crossover_unknowns, char* err, size_t errlen) {
    try {
        const Problem* p = static_cast<Problem*>(h);
        std::optional<Method> m;
        if (method == 0) m = Method::Direct;
        if (method == 1) m = Method::Multigrid;
        Crossover c;
        c.unknowns = crossover_unknowns;
        return new Solver(Solver::make(p->k.submatrix(p->free), p->unknowns, m, c));
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}
// A solver of an assembled matrix (a Csr handle) on grid unknowns given as arrays: the plane
// coordinates of each axis, the grid point (3 plane indices) and the component of each unknown.
void* orc_grid_solver_new(void* csr, const double* px, size_t nx, const double* py, size_t ny, const double* pz,
                          size_t nz, const uint32_t* points, const uint8_t* comp, int method,
                          size_t crossover_unknowns, char* err, size_t errlen) {
    try {
        const Csr& a = *static_cast<Csr*>(csr);
        Unknowns u;
        u.kind = Unknowns::Grid;
        u.level.planes = {std::vector<double>(px, px + nx), std::vector<double>(py, py + ny),
                          std::vector<double>(pz, pz + nz)};
        for (usize i = 0; i < a.n; ++i) {
            u.level.points.push_back({points[3 * i], points[3 * i + 1], points[3 * i + 2]});
            u.level.comp.push_back(comp[i]);
        }
        std::optional<Method> m;
        if (method == 0) m = Method::Direct;
        if (method == 1) m = Method::Multigrid;
        Crossover c;
        c.unknowns = crossover_unknowns;
        return new Solver(Solver::make(a, u, m, c));
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}
// One step of coarsen (multigrid.rs:329-452) on a grid level and its matrix: the prolongation
// (a Rect handle, or null when no axis has more than 2 planes) and the coarse points (3 per coarse
// unknown, written to `coarse_points`, which holds 3 x rows entries at most).
void* orc_coarsen(void* csr, const double* px, size_t nx, const double* py, size_t ny, const double* pz, size_t nz,
                  const uint32_t* points, const uint8_t* comp, uint32_t* coarse_points, uint8_t* coarse_comp) {
    const Csr& a = *static_cast<Csr*>(csr);
    Level level;
    level.planes = {std::vector<double>(px, px + nx), std::vector<double>(py, py + ny), std::vector<double>(pz, pz + nz)};
    for (usize i = 0; i < a.n; ++i) {
        level.points.push_back({points[3 * i], points[3 * i + 1], points[3 * i + 2]});
        level.comp.push_back(comp[i]);
    }
    auto c = coarsen(level, a);
    if (!c) return nullptr;
    for (usize i = 0; i < c->coarse.points.size(); ++i) {
        for (int d = 0; d < 3; ++d) coarse_points[3 * i + d] = c->coarse.points[i][d];
        coarse_comp[i] = c->coarse.comp[i];
    }
    return new Rect(std::move(c->p));
}
int orc_solver_method(void* h) { return (int)static_cast<Solver*>(h)->method(); }
size_t orc_solver_levels(void* h) {
    const Solver* s = static_cast<Solver*>(h);
    if (s->mg) return s->mg->levels();
    if (s->pmg) return 2;
    return 1;
}
size_t orc_solver_level_n(void* h, size_t l) {
    const Solver* s = static_cast<Solver*>(h);
    if (s->mg) return s->mg->a[l].n;
    if (s->pmg) return l == 0 ? s->a.n : s->pmg->coarse_unknowns;
    return s->a.n;
}
void* orc_solver_level_p(void* h, size_t l) {
    const Solver* s = static_cast<Solver*>(h);
    if (s->mg && l < s->mg->p.size()) return new Rect(s->mg->p[l]);
    if (s->pmg && l == 0) return new Rect(s->pmg->p);
    return nullptr;
}
void* orc_solver_level_matrix(void* h, size_t l) {
    const Solver* s = static_cast<Solver*>(h);
    if (s->mg && l < s->mg->a.size()) return new Csr(s->mg->a[l]);
    return nullptr;
}
// Applies the preconditioner once (z = M^-1 r); returns 0, or 1 for the direct solver.
int orc_solver_precondition(void* h, const double* r, double* z) {
    const Solver* s = static_cast<Solver*>(h);
    if (s->mg) {
        s->mg->apply(r, z);
        return 0;
    }
    if (s->pmg) {
        s->pmg->apply(s->a, r, z);
        return 0;
    }
    return 1;
}
int orc_solver_solve(void* h, const double* b, double* x, size_t max_iter, double tol, uint64_t* iterations,
                     double* residual, char* err, size_t errlen) {
    try {
        Solver* s = static_cast<Solver*>(h);
        s->max_iter = max_iter;
        s->tol = tol;
        const Stats st = s->solve(b, x, nullptr);
        *iterations = st.iterations;
        *residual = st.residual;
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}
void orc_solver_free(void* h) { delete static_cast<Solver*>(h); }

}  // extern "C"
