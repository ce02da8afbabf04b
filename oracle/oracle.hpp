// TEST INFRASTRUCTURE ONLY. This is the oracle: a single-threaded FP64 CPU restatement of the
// reference's `fem` solve path (crates/fem/src). Only tests/, the smoke check and the CPU-baseline
// leg of bench.py load it, and only as the checker. The product library never links it.
//
// Numeric conventions follow Rust's: no FMA contraction (built with -ffp-contract=off), no
// fast-math, and every `.sum()` is a sequential left fold, so results are reproducible run to run.
// Bitwise equality with the Rust binary itself cannot be checked here (no Rust toolchain).
//
// Restated files (reference paths relative to the repository root):
//   crates/fem/src/sparse.rs     Csr, Rect, galerkin
//   crates/fem/src/order.rs      nested_dissection
//   crates/fem/src/symbolic.rs   Symbolic::analyse
//   crates/fem/src/cholesky.rs   Cholesky::factor / solve
//   crates/fem/src/multigrid.rs  pcg, Multigrid (coarsen, Lines, cycle)
//   crates/fem/src/pmultigrid.rs PMultigrid, prolongation
//   crates/fem/src/solve.rs      Solver::new / solve
//   crates/fem/src/heat.rs       conduction, tet_conduction, tet_gradients, number, reduced_rhs
//   crates/fem/src/hex.rs        Orthotropic::isotropic / stiffness
//   crates/fem/src/tet10.rs      Tet10, Quadratic
//   crates/fem/src/prism.rs      split_prism and the extrusion's tet orientation rule
//
// How each part is pinned (tests/test_oracle_kats.py):
//   Csr::pattern, Rect::galerkin   the reference's known answers (sparse.rs:300, 308) + scipy P^T A P
//   nested_dissection              order.rs:186 known answer (line of 200: permutation, last 99)
//   Symbolic                       entries == nonzeros of a dense Cholesky factor (grid Laplacian)
//   Cholesky                       cholesky.rs:244 and 268 known answers + scipy spsolve
//   pcg                            solutions equal scipy spsolve; message text of multigrid.rs:730
//   Multigrid / PMultigrid         solutions equal scipy spsolve; every coarse matrix is P^T A P;
//                                  the p-prolongation carries linear fields exactly and P^T K10 P = K4
//   conduction, tet_conduction,    Gauss integration / numpy (V G^T K G) / matrix inverse
//   tet_gradients
//   Orthotropic, Tet10, tet4       hex.rs compliance-inverse test; tet10.rs:300 exact mass; energy of
//                                  uniform strain and of a quadratic field; rigid motions
//   split_prism, Quadratic         tests/tet.rs:92 (720 orders); one node per distinct edge
//   box problems (loads, rhs)      closed forms solved through the direct solver: layered wall
//                                  (tests/heat.rs:112), uniform strain with face tractions, a hanging
//                                  bar under body force (quadratic, exact on tet10)
//   numbering and coarsen          literal vectors written out by hand from the reference's rules
//                                  (tests/test_pinned_numbering.py): the first-seen points and cells
//                                  of a 2 x 2 x 1 hex grid; the points, tetrahedra and Kuhn types of
//                                  1 x 1 x 1 and 2 x 1 x 1 extruded boxes; the edge nodes of a tet10
//                                  cell; coarsen on a two-material 4 x 2 x 2 grid (operator weights
//                                  3/4 and 1/4, geometric 1/2 next to the fixed plane, the first-seen
//                                  coarse numbering with coarse points on that plane)
// Parity unpinned (no reference vector exists and none of the checks above fixes it):
//   - the footprint triangulation of an extruded box: the reference meshes a footprint with its
//     Delaunay mesher (mesh2d.rs), not restated here; the box problems take quads row by row cut
//     along their (0,0)-(1,1) diagonal, and every rule after the footprint is the reference's;
//   - the nested-dissection order on meshes with ties (order.rs:85-101) and the supernode partition
//     beyond the known answers, and therefore the reference's iteration counts;
//   - bitwise equality with the Rust binary.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using u32 = std::uint32_t;
using usize = std::size_t;
using Vec3 = std::array<double, 3>;
constexpr usize NONE = SIZE_MAX;

// A reference `Err(String)`: the text is the reference's text.
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- sparse.rs
// Square CSR with both triangles and sorted columns (sparse.rs:12-17).
struct Csr {
    usize n = 0;
    std::vector<usize> ptr;
    std::vector<u32> idx;
    std::vector<double> val;

    // sparse.rs:22-65. Element k is conn[eptr[k] .. eptr[k+1]].
    static Csr pattern(usize n, const std::vector<usize>& eptr, const std::vector<u32>& conn);
    usize find(usize i, usize j) const;                                // sparse.rs:68
    void add_element(const u32* nodes, usize m, const double* ke);    // sparse.rs:75-86
    std::vector<double> diagonal() const;                              // sparse.rs:89
    void mul(const double* x, double* y) const;                       // sparse.rs:96-104
    Csr submatrix(const std::vector<usize>& keep) const;               // sparse.rs:108-135
};

// Rectangular CSR, the prolongation (sparse.rs:202-208).
struct Rect {
    usize rows = 0, cols = 0;
    std::vector<usize> ptr;
    std::vector<u32> idx;
    std::vector<double> val;

    void mul(const double* x, double* y) const;            // sparse.rs:212
    void mul_transpose(const double* x, double* y) const;  // sparse.rs:223
    Csr galerkin(const Csr& a) const;                      // sparse.rs:233-292
};

// ---------------------------------------------------------------- order.rs
std::vector<usize> nested_dissection(const Csr& a, const std::vector<Vec3>& xyz);  // order.rs:19

// ---------------------------------------------------------------- symbolic.rs
struct Symbolic {
    usize n = 0;
    std::vector<usize> perm, iperm, start, rows_ptr;
    std::vector<u32> rows;
    std::vector<usize> parent, children;

    static Symbolic analyse(const Csr& a, const std::vector<usize>& order);  // symbolic.rs:41-150
    usize supernodes() const { return start.size() - 1; }
    const u32* rows_of(usize s) const { return rows.data() + rows_ptr[s]; }
    usize rows_len(usize s) const { return rows_ptr[s + 1] - rows_ptr[s]; }
    std::uint64_t entries() const;  // symbolic.rs:162
    double flops() const;           // symbolic.rs:174
    std::uint64_t largest_front() const;  // symbolic.rs:185
};

// ---------------------------------------------------------------- cholesky.rs
struct Cholesky {
    Symbolic sym;
    std::vector<usize> offset;
    std::vector<double> values;

    static Cholesky factor(const Symbolic& sym, const Csr& a);  // cholesky.rs:44-122 (throws Error)
    void solve(double* b) const;                                // cholesky.rs:125-160
    usize entries() const { return values.size(); }
};

// ---------------------------------------------------------------- multigrid.rs
struct Level {
    std::array<std::vector<double>, 3> planes;
    std::vector<std::array<u32, 3>> points;
    std::vector<std::uint8_t> comp;
};

struct Convergence {
    usize iterations = 0;
    double residual = 0.0;
};

using Precondition = std::function<void(const double*, double*)>;
using Cancelled = std::function<bool()>;

// multigrid.rs:692-753 (throws Error "conjugate gradients did not converge..." / "cancelled")
Convergence pcg(const Csr& a, const double* b, double* x, const Precondition& precondition, double tol,
                usize max_iter, const Cancelled& cancelled);

struct Lines {
    std::vector<u32> order;
    std::vector<usize> start;
    std::vector<u32> line, place;
    usize band = 0;
    std::vector<double> factor;

    static Lines make(const Level& level, usize axis, const Csr& a);  // multigrid.rs:201-276
    void sweep(const Csr& a, const double* b, double* x, bool forward) const;  // multigrid.rs:280-324
};

struct Coarsened {
    Rect p;
    Level coarse;
    std::array<std::vector<usize>, 3> kept;
};
std::optional<Coarsened> coarsen(const Level& fine, const Csr& a);  // multigrid.rs:329-452

// multigrid.rs:67-70 Cells, with its element closure as the list of element matrices: the lowest
// grid point and the 24 x 24 matrix (unknown 3 n + c, corners in BITS order) of each cell.
struct CellMatrices {
    std::vector<std::array<u32, 3>> origin;
    std::vector<std::array<double, 576>> k;
};

// multigrid.rs:458-600 by_cells: the coarse matrix of a displacement from the cells, and the
// coarse cells and their matrices for the next level. The fine cells of one coarse cell are
// summed in increasing fine-cell order (the reference's unstable sort leaves that order open).
struct ByCells {
    Csr a;
    CellMatrices cells;
};
ByCells by_cells(const Level& fine, const CellMatrices& fc, const Level& coarse, const std::array<std::vector<usize>, 3>& keep);
std::array<bool, 9> coarse_modes(const std::array<std::array<std::array<double, 11>, 11>, 3>& gram);  // multigrid.rs:609-643
std::array<double, 576> condense(std::vector<double>& k, const std::array<bool, 9>& keep);            // multigrid.rs:648-671

struct Multigrid {
    std::vector<Csr> a;
    std::vector<Rect> p;
    std::vector<std::array<Lines, 3>> lines;
    Cholesky coarse;

    // multigrid.rs:84-139: Galerkin coarse matrices (cells = None), or with `cells` (a
    // displacement) the coarse matrices of by_cells.
    static Multigrid make(const Csr& a, Level level, const CellMatrices* cells = nullptr);
    usize levels() const { return a.size(); }
    void apply(const double* r, double* z) const;  // multigrid.rs:147
    void cycle(usize l, const double* b, double* x) const;  // multigrid.rs:151-179
};

// ---------------------------------------------------------------- pmultigrid.rs
struct PMultigrid {
    Rect p;
    std::vector<double> diag;
    Cholesky coarse;
    usize coarse_unknowns = 0;

    static PMultigrid make(const Csr& a, Rect p, const std::vector<Vec3>& coarse);  // pmultigrid.rs:46-60
    void apply(const Csr& a, const double* r, double* z) const;  // pmultigrid.rs:63-98
};
// pmultigrid.rs:104-151
std::pair<Rect, std::vector<Vec3>> prolongation(const std::vector<Vec3>& points, usize corners,
                                                const std::vector<std::array<u32, 2>>& edges,
                                                const std::vector<usize>& free);

// ---------------------------------------------------------------- solve.rs
constexpr usize STEADY_DIRECT_LIMIT = 3000;     // solve.rs:18
constexpr usize TRANSIENT_DIRECT_LIMIT = 40000; // solve.rs:24
constexpr double PCG_TOL = 1e-10;               // solve.rs:42
constexpr usize PCG_MAX_ITER = 500;             // solve.rs:43

enum class Method { Direct = 0, Multigrid = 1 };

struct Crossover {  // solve.rs:28-39
    bool by_flops = false;
    usize unknowns = STEADY_DIRECT_LIMIT;
    double flops = 0.0;
    std::uint64_t entries = 0;
    bool flat = false;
};

struct Unknowns {  // solve.rs:64-77
    enum Kind { Grid, Points, Quadratic } kind = Grid;
    Level level;                 // Grid
    std::vector<Vec3> points;    // Points, Quadratic
    Rect p;                      // Quadratic
    std::vector<Vec3> coarse;    // Quadratic
};

struct Stats {  // solve.rs:81-86
    usize iterations = 0;
    double residual = 0.0;
};

struct Solver {
    enum Inner { Direct, MG, PMG } inner = Direct;
    Csr a;
    std::optional<Cholesky> chol;
    std::optional<Multigrid> mg;
    std::optional<PMultigrid> pmg;
    usize max_iter = PCG_MAX_ITER;
    double tol = PCG_TOL;  // the reference fixes PCG_TOL; the oracle lets a test tighten it

    static Solver make(Csr a, Unknowns unknowns, std::optional<Method> method, Crossover crossover,
                       const CellMatrices* cells = nullptr);
    Method method() const { return inner == Direct ? Method::Direct : Method::Multigrid; }
    Stats solve(const double* b, double* x, const Cancelled& cancelled) const;  // solve.rs:199-250
};

// ---------------------------------------------------------------- elements
std::array<double, 64> conduction(Vec3 s, Vec3 k);                        // heat.rs:1246-1278
std::array<double, 16> tet_conduction(const std::array<Vec3, 4>& g, double v, Vec3 k);  // heat.rs:972
std::pair<std::array<Vec3, 4>, double> tet_gradients(const std::array<Vec3, 4>& p);     // heat.rs:983

struct Stiffness {  // hex.rs:120-124
    double normal[3][3];
    double shear[3];  // G_yz, G_xz, G_xy
};
struct Orthotropic {  // hex.rs:47-55
    Vec3 e, nu, g;
    static Orthotropic isotropic(double e, double nu);  // hex.rs:58
    Stiffness stiffness() const;                        // hex.rs:69-115 (throws Error)
};

extern const usize TET10_EDGES[6][2];  // tet10.rs:38
struct Tet10 {  // tet10.rs:64-213
    std::array<Vec3, 4> g;
    double volume = 0.0;
    static Tet10 make(const std::array<Vec3, 4>& p);
    std::array<Vec3, 10> gradients(const std::array<double, 4>& l) const;
    std::vector<double> stiffness(const Stiffness& d) const;  // 30 x 30 row-major
    std::vector<double> mass(double density) const;
};
std::array<std::array<double, 4>, 4> tet10_gauss();  // tet10.rs:44-48

// The constant-strain tetrahedron of the same material law: V B^T D B, unknown 3 i + a. The
// reference forms this operator only as the Galerkin product P^T K10 P of the p-multigrid
// (pmultigrid.rs:7-8); a test checks the two agree.
std::vector<double> tet4_stiffness(const std::array<Vec3, 4>& g, double volume, const Stiffness& d);

std::array<std::array<usize, 4>, 3> split_prism(const std::array<std::uint64_t, 6>& g);  // prism.rs:144

std::array<double, 8> hex_shape(const Vec3& xi);                              // hex.rs:164-178
std::array<double, 576> hex_stiffness(const Vec3& size, const Stiffness& d);  // hex.rs:199-279 Hex::new (k)

struct Quadratic {  // tet10.rs:232-272
    std::vector<Vec3> points;
    usize corners = 0;
    std::vector<std::array<u32, 2>> edges;
    std::vector<std::array<u32, 10>> tets;
    static Quadratic make(const std::vector<Vec3>& points, const std::vector<std::array<u32, 4>>& tets);
};

// ---------------------------------------------------------------- structured box problems
// The reference callers (heat.rs:755-877 + 503 + 661-673; elastic_tet.rs:88-331) assembled on a
// structured box so that both the oracle and the GPU library build the same model.

// A box of nx x ny x nz cells, [0,lx] x [0,ly] x [0,lz].
struct BoxSpec {
    int nx = 1, ny = 1, nz = 1;
    double lx = 1, ly = 1, lz = 1;
};

// Faces: 0 x=0, 1 x=L, 2 y=0, 3 y=L, 4 z=0, 5 z=L.
struct HeatBoxSpec {
    BoxSpec box;
    std::vector<double> k;           // per cell (hex8) or per tet (tet4): isotropic conductivity
    std::array<bool, 6> fixed{};     // a Dirichlet face
    std::array<double, 6> value{};   // its temperature
    double source = 0.0;             // uniform heat source per volume
};

struct ElasticBoxSpec {
    BoxSpec box;
    int order = 10;                  // 4 or 10
    std::vector<double> e, nu;       // per tet
    std::array<bool, 6> clamped{};   // all 3 components held at 0
    Vec3 body{0, 0, 0};              // force per volume
    std::array<Vec3, 6> traction{};  // force per area on each face
};

// Box-grid linear elasticity on hex8 cells (elastic.rs:287-563): E and nu per cell.
struct ElasticHexSpec {
    BoxSpec box;
    std::vector<double> e, nu;       // per cell
    std::array<bool, 6> clamped{};
    Vec3 body{0, 0, 0};              // force per volume, V/8 on each corner (elastic.rs:402)
};

struct Problem {
    usize n_points = 0;
    usize comps = 1;
    std::vector<Vec3> points;        // coordinates, reference numbering
    std::vector<std::array<u32, 3>> ijk;  // hex8: grid point; tets: refined-grid point (2x for tet10)
    std::vector<u32> cells;          // hex8: 8 per cell; tet4: 4 per tet; tet10: 10 per tet
    usize npe = 0;
    Csr k;                           // over all unknowns
    std::vector<double> load;        // over all unknowns
    std::vector<double> fixed_value; // over all unknowns (0 if free)
    std::vector<std::uint8_t> is_fixed;
    std::vector<usize> free;         // increasing
    std::vector<double> rhs;         // reduced right side over free unknowns
    Unknowns unknowns;
    CellMatrices cell_matrices;      // hex elasticity: each cell's lowest grid point and element matrix
    // tet meshes
    std::vector<std::array<u32, 4>> tets4;
    usize corners = 0;
    std::vector<std::array<u32, 2>> edges;
};

Problem heat_hex8_box(const HeatBoxSpec& spec);
Problem heat_tet4_box(const HeatBoxSpec& spec);
Problem elastic_tet_box(const ElasticBoxSpec& spec);
Problem elastic_hex8_box(const ElasticHexSpec& spec);

// The tet4 corner mesh of a box: the footprint (x, y) of quads cut at their lowest-corner
// diagonal, extruded along z, each prism split by split_prism, numbered first-seen
// (prism.rs:244-284). `ijk` is the grid point of each corner.
void box_tets(const BoxSpec& box, std::vector<Vec3>& points, std::vector<std::array<u32, 3>>& ijk,
              std::vector<std::array<u32, 4>>& tets);

// heat.rs:1219-1242 (steady: no capacity term)
std::vector<double> reduced_rhs(const Csr& a, const std::vector<double>& load, const std::vector<double>& t,
                                const std::vector<usize>& free);

}  // namespace orc
