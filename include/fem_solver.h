/*
 * femahoy_amd: the iterative solve of femahoy's `fem` crate (PCG preconditioned by a geometric /
 * p-multigrid V-cycle, for steady heat conduction and linear elasticity) on one AMD MI355X.
 *
 * A thin C ABI: plain pointers and sizes, no C++ or torch types. Each entry point names the
 * reference interface it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *  - Return 0 on success. A nonzero return leaves a message in `err` (up to `errlen` bytes, NUL
 *    terminated). Messages that have a counterpart in the reference use the reference's text.
 *  - A model is a structured box: nx x ny x nz cells on [0,lx] x [0,ly] x [0,lz]. Heat on hex8
 *    cells or on tet4 (each cell cut in the 6 tetrahedra of the reference's prism split), or
 *    linear elasticity on tet4 / tet10. Unknowns are numbered as the reference numbers them
 *    (crates/fem/src/heat.rs:1165-1208 for grid points, prism.rs:244-284 and tet10.rs:244-272 for
 *    tetrahedra, unknown 3 node + component for a displacement, elastic_tet.rs:178-183), and the
 *    vectors of a solve hold the free unknowns in increasing order (heat.rs:503,
 *    elastic_tet.rs:291).
 *  - The caller keeps its host arrays; the library copies what it needs into device memory it
 *    owns. A solver handle is not safe for concurrent calls; distinct handles are.
 */
#ifndef FEM_SOLVER_H
#define FEM_SOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The element and physics of a box model. */
enum fem_kind {
    FEM_HEAT_HEX8 = 0,     /* heat.rs:1246 conduction() on box cells */
    FEM_HEAT_TET4 = 1,     /* heat.rs:972 tet_conduction() */
    FEM_ELASTIC_TET4 = 2,  /* constant-strain tetrahedron (the p-multigrid coarse operator, pmultigrid.rs:7) */
    FEM_ELASTIC_TET10 = 3, /* tet10.rs:90 Tet10::stiffness() */
    FEM_ELASTIC_HEX8 = 4   /* hex.rs:199-279 Hex::new: the box-grid elasticity caller (elastic.rs:287-547), assembled */
};

/* solve.rs:47-50 Method, plus "let the crossover choose". */
enum fem_method { FEM_METHOD_AUTO = -1, FEM_METHOD_DIRECT = 0, FEM_METHOD_MULTIGRID = 1 };

/* solve.rs:28-39 Crossover: how FEM_METHOD_AUTO chooses. UNKNOWNS: direct up to `crossover`
 * unknowns. FLOPS: direct when the factor of the nested dissection order (order.rs:19-53,
 * symbolic.rs:41-192) has at most `crossover_entries` entries, and at most `crossover_flops` flops
 * per unknown or `crossover_flat` (solve.rs:106-118, 158-170). */
enum fem_crossover { FEM_CROSSOVER_UNKNOWNS = 0, FEM_CROSSOVER_FLOPS = 1 };

/* The arithmetic of the preconditioner. The outer PCG (residual recurrence, dots, search
 * direction, operator of the outer iteration, true residual) is FP64 in every mode. */
enum fem_precision {
    FEM_PREC_FP64 = 0,  /* FP64 V-cycle */
    FEM_PREC_FP32 = 1,  /* FP32 V-cycle (storage and arithmetic) */
    FEM_PREC_BF16 = 2   /* FP32 V-cycle with BF16-stored coefficients, FP32 accumulation */
};

/* A structured box model. Faces are numbered 0: x=0, 1: x=lx, 2: y=0, 3: y=ly, 4: z=0, 5: z=lz. */
typedef struct fem_box {
    int32_t kind;         /* enum fem_kind */
    int32_t n[3];         /* cells along x, y, z */
    double length[3];     /* box size */
    const double* coef0;  /* heat: conductivity of each cell (hex8) or tetrahedron (tet4);
                             elasticity: Young's modulus of each tetrahedron (tet4, tet10) or cell
                             (hex8). Reference order. */
    const double* coef1;  /* elasticity: Poisson ratio of each tetrahedron or cell; heat: unused */
    int32_t fixed[6];     /* heat: a face at a fixed temperature; elasticity: a clamped face */
    double value[6];      /* heat: the temperature of a fixed face */
    double source;        /* heat: heat source per volume */
    double body[3];       /* elasticity: force per volume */
    double traction[18];  /* elasticity: force per area on each face, 3 per face (hex8: a quarter of
                             each cell face's force on each of its 4 points, elastic.rs:942-963) */
} fem_box;

/* Options of a solver. `fem_options_default` fills the reference's values where it has them. */
typedef struct fem_options {
    double tol;             /* relative recurrence residual; solve.rs:42 PCG_TOL = 1e-10 */
    uint64_t max_iter;      /* solve.rs:43 PCG_MAX_ITER = 500; Solver::with_max_iter (solve.rs:187) */
    int32_t method;         /* enum fem_method */
    int32_t precision;      /* enum fem_precision */
    uint64_t crossover;     /* direct up to this many unknowns with FEM_METHOD_AUTO; solve.rs:18 (3000) */
    int32_t levels;         /* V-cycle levels including the finest; 0 = automatic: down to max(1000, free/256) unknowns, at most coarse_max */
    int32_t cheb_degree;    /* Chebyshev-Jacobi smoother degree (pre and post) on every level; 0 =
                               automatic: for a box, 3 on the finest level and 1 below for scalar
                               problems, 2 on every level for vector ones; 2 for fem_solver_new_csr */
    double cheb_lower;      /* smoothed range [cheb_lower * lmax, cheb_upper * lmax] of D^-1 A */
    double cheb_upper;
    int32_t power_steps;    /* power iterations of the lmax estimate on each level */
    uint64_t coarse_max;    /* largest dense coarse level (unknowns) */
    int32_t refine;         /* 1: after convergence of the recurrence, recompute the FP64 true
                               residual and restart PCG from x while it is above tol (at most 5) */
    int32_t device;         /* HIP device ordinal */
    int32_t coarse_galerkin; /* hex8: 1 Galerkin coarse operators (27-point stencils, as the reference's
                                P^T A P, multigrid.rs:109); 0 the hex8 operator rediscretized with the
                                mean conductivity of the 8 children */
    int32_t crossover_kind;     /* enum fem_crossover (default UNKNOWNS) */
    double crossover_flops;     /* FLOPS: the most factor flops per unknown of a direct solve
                                   (elastic.rs:263 DIRECT_FLOPS = 1.5e5; infinite: the entries decide) */
    uint64_t crossover_entries; /* FLOPS: the most factor entries (elastic.rs:269 DIRECT_ENTRIES = 2.5e8) */
    int32_t crossover_flat;     /* FLOPS: 1 when the cells are flat (direct up to crossover_entries) */
    int32_t direct_solver;      /* enum fem_direct: the factor of the direct method of fem_solver_new_csr */
} fem_options;

/* The factor of the direct method (fem_solver_new_csr). The reference's is its multifrontal sparse
 * Cholesky (cholesky.rs:44-160) in the nested-dissection order (order.rs:19-53). AUTO: the dense
 * LDL^T up to 8000 unknowns, the multifrontal LDL^T above (and where the unknowns have no
 * coordinates, the dense one up to 40000). */
enum fem_direct { FEM_DIRECT_AUTO = 0, FEM_DIRECT_DENSE = 1, FEM_DIRECT_MULTIFRONTAL = 2 };

/* solve.rs:81-86 Stats first ({iterations, residual}: the recurrence residual of PCG, the true
 * residual of the direct solver), then what the GPU solve adds. */
typedef struct fem_stats {
    uint64_t iterations;
    double residual;
    double true_residual;    /* |b - A x| / |b| in FP64 at the end */
    double seconds;          /* device time of the solve: right side resident to solution resident */
    double setup_seconds;    /* hierarchy, eigenvalue estimates, coarse factor */
    uint64_t levels;
    uint64_t coarse_unknowns;
    uint64_t restarts;       /* refinement restarts */
    int32_t method;          /* enum fem_method (never AUTO) */
    int32_t precision;       /* enum fem_precision */
    uint64_t hbm_bytes;      /* algorithmic bytes of the solve's launches: every array a launch reads or
                                writes, once per launch at its stored precision (box solvers; 0 for
                                the assembled boundary and the direct method) */
} fem_stats;

typedef struct fem_solver fem_solver;

/* Fills the defaults: tol 1e-10, max_iter 500, AUTO, FP32 V-cycle, crossover UNKNOWNS 3000 (with
 * the FLOPS values of elastic.rs: 1.5e5 flops per unknown, 2.5e8 entries, not flat), automatic
 * levels, automatic Chebyshev degree on [0.1, 1.1] lmax with 20 power steps, coarse_max 6000,
 * refine 1, Galerkin coarse operators. */
void fem_options_default(fem_options* opt);

/* Replaces Solver::new (crates/fem/src/solve.rs:92-183) for a box model: builds the model, its
 * numbering and free unknowns, and then the direct factor (a dense FP64 LDL^T on the device,
 * up to 40000 unknowns) or the multigrid hierarchy. With FEM_METHOD_AUTO the crossover chooses as
 * the reference's does (solve.rs:106-118, 158-170): Crossover::Unknowns up to `crossover`
 * unknowns, or Crossover::Flops on the counts of the nested-dissection factor of the model's
 * reduced matrix pattern. Unlike the reference, multigrid also runs on tetrahedra (solve.rs:132-137
 * refuses it there). Material errors use the texts of hex.rs:79 and hex.rs:93.
 * FEM_ELASTIC_HEX8 is solved as the reference's box-grid elasticity caller solves it
 * (elastic.rs:287-547): the condensed incompatible-mode hex of each cell (hex.rs:199-279) is
 * assembled on the host into the reduced matrix of the free unknowns, which goes with the grid
 * level of its unknowns and the cells (for the coarse matrices of multigrid.rs:458-671) to the
 * solver of fem_solver_new_csr; the caller chooses the crossover (the reference's caller passes
 * Crossover::Flops with `flat` when a cell's sides differ by more than 2x, elastic.rs:525-547).
 * fem_solver_rhs, _free_list, _point_grid and _matrix give its right side, free unknowns, points
 * and reduced matrix. The serving loop takes the matrix-free kinds only. */
int fem_solver_new_box(const fem_box* box, const fem_options* opt, fem_solver** out, char* err, size_t errlen);

/* Replaces Solver::solve (solve.rs:199-250) with host vectors: `b` and `x` hold the free
 * unknowns in the reference order; `x` is the start value of PCG and gets the solution.
 * `cancelled` (may be NULL) is polled once per PCG iteration, as the reference's closure
 * (multigrid.rs:735-737). Errors: "conjugate gradients did not converge: the residual is {:.2e}
 * after {} iterations" (multigrid.rs:730-733) and "cancelled" (multigrid.rs:736). */
int fem_solver_solve(fem_solver* s, const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* stats,
                     char* err, size_t errlen);

/* The same solve with `b` and `x` already in device memory (HIP pointers, free unknowns in the
 * reference order): no host transfer inside. */
int fem_solver_solve_device(fem_solver* s, const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx,
                            fem_stats* stats, char* err, size_t errlen);

/* Several right sides against one setup (no single counterpart in the reference; it replaces the
 * loop of Solver::solve calls on one Solver of its transient steps, heat.rs:697-735, DC terminals,
 * dc.rs:262-271, and shift-invert blocks, modes.rs:412). b and x hold k vectors of the free
 * unknowns one after another (b + i n, x + i n); each x_i is that solve's start value and gets its
 * solution; stats holds k entries. A box multigrid solver runs the k PCG solves together: it keeps
 * k - 1 shadow solvers that share its hierarchy (coefficients, D^-1, coarse inverse) and own only
 * their vectors, and batches their iterations (one launch per kernel over the same level of every
 * right side). Each result equals fem_solver_solve of that right side, bit for bit. Errors as
 * fem_solver_solve; after all k have ended, the first error is returned. */
int fem_solver_solve_many(fem_solver* s, uint64_t k, const double* b, double* x, int (*cancelled)(void*), void* ctx,
                          fem_stats* stats, char* err, size_t errlen);
/* The same with b and x in device memory (k n values each). */
int fem_solver_solve_many_device(fem_solver* s, uint64_t k, const double* d_b, double* d_x, int (*cancelled)(void*),
                                 void* ctx, fem_stats* stats, char* err, size_t errlen);

/* Solver::method (solve.rs:191). */
int fem_solver_method(const fem_solver* s);

/* Solver::with_max_iter (solve.rs:187), in place. */
void fem_solver_set_max_iter(fem_solver* s, uint64_t max_iter);

/* Sets the relative tolerance of later solves (the reference fixes PCG_TOL). */
void fem_solver_set_tol(fem_solver* s, double tol);

void fem_solver_free(fem_solver* s);

/* ---- the assembled boundary: a CSR matrix and its unknowns, as Solver::new takes them ---- */

/* Where the unknowns are (solve.rs:64-77 Unknowns). */
enum fem_unknowns_kind {
    FEM_UNKNOWNS_GRID = 0,     /* Unknowns::Grid(Level): grid multigrid (multigrid.rs:84-139) or direct */
    FEM_UNKNOWNS_POINTS = 1,   /* Unknowns::Points: direct only (solve.rs:132-137) */
    FEM_UNKNOWNS_QUADRATIC = 2 /* Unknowns::Quadratic: p-multigrid (pmultigrid.rs:46-60) or direct */
};

/* A sparse matrix in CSR form (sparse.rs:12-17, and Rect, sparse.rs:202-208): row i holds the
 * columns idx[ptr[i] .. ptr[i+1]), increasing, with the values val at the same places. */
typedef struct fem_csr {
    uint64_t rows, cols;   /* a square matrix has cols == rows */
    const uint64_t* ptr;   /* rows + 1 offsets */
    const uint32_t* idx;   /* ptr[rows] columns */
    const double* val;     /* ptr[rows] values */
} fem_csr;

/* A grid level (multigrid.rs:55-59 Level): the plane coordinates of each axis, increasing, and the
 * grid point (3 plane indices) and the component (0 for a scalar, 0..2 for a displacement) of
 * each unknown. */
typedef struct fem_grid_level {
    const double* planes[3];
    uint64_t nplanes[3];
    const uint32_t* points;  /* 3 per unknown */
    const uint8_t* comp;     /* 1 per unknown */
} fem_grid_level;

/* multigrid.rs:67-70 Cells: the cells of the fine level of a displacement model, the lowest grid
 * point of each cell and its 24 x 24 element matrix (row-major, unknown 3 n + c for corner n in the
 * order of hex.rs:34 BITS and component c), from which the reference builds the coarse matrices
 * (multigrid.rs:458-679 by_cells). */
typedef struct fem_cells {
    uint64_t count;
    const uint32_t* origin;  /* 3 per cell */
    const double* element;   /* 576 per cell */
} fem_cells;

typedef struct fem_unknowns {
    int32_t kind;              /* enum fem_unknowns_kind */
    fem_grid_level grid;       /* GRID: the level of the matrix's unknowns */
    fem_csr p;                 /* QUADRATIC: the prolongation from the linear unknowns (rows = unknowns,
                                  cols = linear unknowns; pmultigrid.rs:104-151 makes it) */
    int32_t coarse_has_grid;   /* QUADRATIC: 1 when coarse_grid gives the linear unknowns' grid level, so
                                  that grid multigrid continues below the linear level (not in the
                                  reference, which factors that level: pmultigrid.rs:46-60) */
    fem_grid_level coarse_grid;
    const double* xyz;         /* POINTS and QUADRATIC: the coordinates of each unknown (3 per unknown;
                                  Unknowns::Points(xyz), Unknowns::Quadratic { points }), read by the
                                  FLOPS crossover's nested dissection; may be NULL otherwise (GRID
                                  unknowns take theirs from the grid level, solve.rs:143-150) */
    const fem_cells* cells;    /* Solver::new's `cells` (solve.rs:97): NULL for None */
} fem_unknowns;

/* Replaces Solver::new(a, unknowns, method, crossover, cells) (crates/fem/src/solve.rs:92-183)
 * for an assembled matrix of the free unknowns (symmetric positive definite, both triangles).
 * The method follows the reference: direct with FEM_METHOD_DIRECT; with FEM_METHOD_AUTO as the
 * crossover chooses (Crossover::Unknowns up to `crossover` unknowns, or Crossover::Flops from the
 * nested dissection factor's counts, which need the unknowns' coordinates: the grid level's, or
 * `xyz`); else grid multigrid (GRID) or p-multigrid (QUADRATIC); POINTS is always direct, and
 * POINTS with FEM_METHOD_MULTIGRID fails with the text of solve.rs:135. The grid hierarchy is the
 * reference's (multigrid.rs:84-139: coarsen, 329-452, and Galerkin products, sparse.rs:233-292,
 * down to 3000 unknowns; `levels` > 0 stops earlier); its smoother is Chebyshev-Jacobi and its
 * coarsest level the dense LDL^T (up to coarse_max unknowns; a QUADRATIC linear level with no grid
 * below it up to 40000). The direct solver (fem_options.direct_solver) is the dense LDL^T up to
 * 8000 unknowns and the multifrontal LDL^T above: the nested-dissection order and fundamental
 * supernodes of the reference's factor (order.rs:19-53, symbolic.rs:41-150, cholesky.rs:44-160),
 * its fronts factored on the device level by level of the supernode tree, its pivot failure in the
 * reference's text (cholesky.rs:92-99); it needs the unknowns' coordinates (GRID, or `xyz`), and
 * without them the dense LDL^T takes up to 40000 unknowns and refuses more with the factor's counts.
 * With `cells` (GRID unknowns of a
 * displacement) the coarse matrices are those of by_cells (multigrid.rs:458-671: each coarse cell's
 * energy on its trilinear corner fields and incompatible modes, condensed), as the reference's
 * box-grid elasticity caller asks (elastic.rs:537-547). b and x of fem_solver_solve are the matrix's
 * unknowns in its order; fem_solver_rhs and fem_solver_point_grid do not apply. */
int fem_solver_new_csr(const fem_csr* a, const fem_unknowns* u, const fem_options* opt, fem_solver** out, char* err,
                       size_t errlen);

/* ---- assembly on the device: a mesh to the reduced matrix and right side ---- */

/* The element kinds of an assembled mesh. */
enum fem_mesh_kind {
    FEM_MESH_HEAT_HEX8 = 0,     /* conduction of axis-aligned box cells (heat.rs:1246-1278), 8 corners
                                   in VTK order (hex.rs:34-43); cell sizes from corners 0 and 6 */
    FEM_MESH_HEAT_TET4 = 1,     /* conduction of straight tetrahedra (heat.rs:972-1000) */
    FEM_MESH_ELASTIC_TET10 = 2, /* isotropic elasticity of straight tet10 (tet10.rs:64-120): 4 corners, then
                                   the 6 edge nodes in the order of tet10.rs:38 */
    FEM_MESH_ELASTIC_HEX8 = 3   /* isotropic elasticity of axis-aligned box cells with incompatible modes
                                   (hex.rs:199-279, the box-grid caller elastic.rs:287-547), 8 corners in
                                   VTK order, sizes from corners 0 and 6; the element of each distinct
                                   (sizes, material) on the host, the rest on the device */
};

/* A mesh and its loads and held unknowns, as the reference's callers hold them before
 * Csr::pattern / add_element / submatrix / reduced_rhs (sparse.rs:22-86; heat.rs:661-673 and
 * 1219-1242; elastic_tet.rs:178-331). Unknowns: one per node (heat), 3 node + component
 * (elasticity, elastic_tet.rs:178-183). */
typedef struct fem_mesh {
    int32_t kind;            /* enum fem_mesh_kind */
    uint64_t points;
    const double* xyz;       /* 3 per point */
    uint64_t cells;
    const uint32_t* conn;    /* 8, 4 or 10 points per cell */
    const double* coef0;     /* per cell: the isotropic conductivity, or Young's modulus */
    const double* coef1;     /* per cell: Poisson's ratio (elasticity); may be NULL for heat */
    double source;           /* heat: source per volume, V/8 (hex8) or V/4 (tet4) to each corner */
    double body[3];          /* elasticity: force per volume; tet10: -V/20 to each corner and V/5 to each
                                edge node; hex8: V/8 to each corner (elastic.rs:384-392) */
    const uint8_t* fixed;    /* per unknown: 1 when held */
    const double* value;     /* per unknown: its held value (read where fixed) */
    const double* load;      /* per unknown: a further load (face loads) added after the volume loads; may be NULL */
    const double* ortho;     /* elasticity: per cell 9 orthotropic constants (E_x, E_y, E_z, nu_xy, nu_xz, nu_yz,
                                G_xy, G_xz, G_yz; hex.rs:47-115) in place of coef0 and coef1; may be NULL */
} fem_mesh;

/* The assembled system of a mesh on the device (no counterpart type in the reference). */
typedef struct fem_assembly fem_assembly;

/* Timings of an assembly. */
typedef struct fem_assembly_stats {
    uint64_t free;           /* unknowns of the reduced system */
    uint64_t entries;        /* its stored entries (both triangles) */
    uint64_t full_entries;   /* the entries of the whole pattern */
    double host_seconds;     /* validation, the pattern, the free list and the gather maps (host) */
    double device_seconds;   /* element matrices and loads, the gathers, the reduced matrix and right side */
} fem_assembly_stats;

/* Assembles a mesh on the device: the pattern of Csr::pattern (sparse.rs:22-62) and the free
 * unknowns (increasing) on the host, then on the device every element's matrix and loads (one
 * thread a matrix row), each matrix entry and load summed from its elements in element order as
 * add_element and the callers' load loops sum them (sparse.rs:64-86), the reduced matrix
 * (Csr::submatrix over the free unknowns) and the reduced right side of heat.rs:1219-1242. FP64
 * throughout, with no contraction of products into fused multiply-adds, so the values equal a
 * sequential CPU assembly in the reference's order bit for bit. A material outside the
 * reference's range gives its text (hex.rs:79 and 93). The call waits for the device. */
int fem_assemble(const fem_mesh* mesh, int device, fem_assembly** out, char* err, size_t errlen);
void fem_assembly_stats_of(const fem_assembly* a, fem_assembly_stats* out);
/* The reduced system to host arrays: ptr (free + 1), idx and val (entries), rhs (free), and the
 * free unknowns (free); any may be NULL. */
int fem_assembly_download(const fem_assembly* a, uint64_t* ptr, uint32_t* idx, double* val, double* rhs,
                          int64_t* free_list, char* err, size_t errlen);
/* The device arrays of the reduced system, valid while the assembly lives. */
void fem_assembly_device(const fem_assembly* a, const uint64_t** ptr, const uint32_t** idx, const double** val,
                         const double** rhs);
void fem_assembly_free(fem_assembly* a);

/* The condensed incompatible-mode hex of an axis-aligned box cell with sides size[3] (hex.rs:199-279
 * Hex::new) of an orthotropic material with its axes on x, y and z (hex.rs:47-115 Orthotropic):
 * material[9] = E_x, E_y, E_z, nu_xy, nu_xz, nu_yz, G_xy, G_xz, G_yz (an isotropic material: E, nu
 * and E / (2 (1 + nu)) on every axis). Its stiffness k (24 x 24) and, for the expansion
 * coefficients alpha[3] along x, y and z (may be NULL: none), the load of a unit temperature change
 * at each corner (24 x 8; may be NULL), row-major, unknown 3 corner + component, corners in VTK
 * order. The element that FEM_ELASTIC_HEX8 boxes and FEM_MESH_ELASTIC_HEX8 meshes assemble, for
 * callers that build their own cells and thermal loads (the case-file front end). Material errors
 * use the texts of hex.rs:79 and hex.rs:93. */
int fem_hex_element(const double* size, const double* material, const double* alpha, double* k, double* load,
                    char* err, size_t errlen);

/* hex.rs:313-340 Hex::strain_stress of the same element: the engineering strain (xx, yy, zz, yz,
 * xz, xy) and the stress in the same order at the local point xi[3] (each coordinate -1 to 1) of
 * the cell, from its corner displacements u[24] (3 corner + component) and corner temperature
 * changes t[8]: the incompatible modes' amplitudes recovered from u and t, the strain of the 11
 * functions, D times it less the thermal stress of the interpolated temperature change. Host
 * code. The probes' stresses of the case-file front end (elastic.rs:676-700). */
int fem_hex_stress(const double* size, const double* material, const double* alpha, const double* xi,
                   const double* u, const double* t, double* strain, double* stress, char* err, size_t errlen);

/* ---- a dense direct solve on device memory ---- */

/* x = A^-1 b for a dense symmetric positive definite n x n matrix A: the device's dense FP64
 * LDL^T (blocked, on the FP64 matrix cores), its explicit inverse, and one step of refinement on
 * the FP64 residual, the same factor and solve as the direct method of fem_solver_new_box and
 * fem_solver_new_csr. It replaces Cholesky::factor and Cholesky::solve
 * (crates/fem/src/cholesky.rs:44-160) for a matrix held dense. d_a (row-major, n x n, FP64),
 * d_b and d_x (n) are device pointers; d_a and d_b are not changed. The device workspace (three
 * n^2 arrays) is allocated for the call and freed before it returns; the call waits for the device.
 * A failed pivot gives the reference's "the matrix is not positive definite: ..." text
 * (cholesky.rs:94-98). */
int fem_dense_solve_device(int n, const double* d_a, const double* d_b, double* d_x, char* err, size_t errlen);

/* The same solve with no device allocation: no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its workspace is one fixed array of device memory (25.1 MiB)
 * that each device the library runs on holds until the process ends, so it takes matrices of at
 * most 1024 unknowns and refuses larger ones with a message. Calls on one device share that
 * array: calls from several host threads of a process on one device run one after another (a
 * lock per device), as they would otherwise race on it. Results equal fem_dense_solve_device's
 * bit for bit. The call waits for the device. */
int fem_dense_solve_device_noalloc(int n, const double* d_a, const double* d_b, double* d_x, char* err,
                                   size_t errlen);

/* The counts of the nested-dissection factor that a FLOPS crossover compared (symbolic.rs:162-183
 * entries() and flops()), when the solver's choice read them. Returns 0 and the counts, or 1 when
 * the crossover did not analyse the matrix (UNKNOWNS, or a forced method). */
int fem_solver_crossover_counts(const fem_solver* s, uint64_t* entries, double* flops);

/* ---- the model's integer structures and right side, for the caller and for parity checks ---- */

/* The number of free unknowns (the length of b and x), and of all unknowns. */
uint64_t fem_solver_free_count(const fem_solver* s);
uint64_t fem_solver_unknown_count(const fem_solver* s);

/* The free unknowns in increasing order (heat.rs:503, elastic_tet.rs:291): `out` holds
 * fem_solver_free_count values. */
int fem_solver_free_list(const fem_solver* s, int64_t* out);

/* The structured grid point of each point in the reference numbering (3 per point; for tet10
 * the point on the grid refined twice, corners at even coordinates). `out` holds 3 x points. */
int fem_solver_point_grid(const fem_solver* s, uint32_t* out);

/* The reduced right side of the model (heat.rs:1219-1242 reduced_rhs: load minus the fixed
 * values' coupling; elastic_tet.rs:202-266 loads), over the free unknowns in reference order. */
int fem_solver_rhs(fem_solver* s, double* out, char* err, size_t errlen);

/* y = A x over the free unknowns (reference order, FP64, host vectors): the reduced matrix of
 * Csr::submatrix(free) (sparse.rs:108) times x, as Csr::mul (sparse.rs:96). */
int fem_solver_apply(fem_solver* s, const double* x, double* y, char* err, size_t errlen);

/* The reduced matrix of a box model, as the reference's callers assemble it before Solver::new:
 * Csr::submatrix(free) (sparse.rs:108-135) of the Csr::pattern of the elements (sparse.rs:22-65),
 * over the free unknowns in reference order, columns increasing, FP64 values of the device's
 * operator. Call with ptr NULL for the entry count (*nnz); then with ptr (free count + 1), idx and
 * val (nnz each). hex8, tet4 and tet10 models. */
int fem_solver_matrix(fem_solver* s, uint64_t* nnz, uint64_t* ptr, uint32_t* idx, double* val, char* err,
                      size_t errlen);

/* Applies the V-cycle preconditioner once, z = M^-1 r (host vectors, reference order). */
int fem_solver_precondition(fem_solver* s, const double* r, double* z, char* err, size_t errlen);

/* Hierarchy facts: the free unknowns of level l, and the Chebyshev bound lmax used there. */
int fem_solver_level_info(const fem_solver* s, int32_t level, uint64_t* unknowns, double* lmax);

/* The name of the HIP device the solver runs on. */
const char* fem_solver_device_name(const fem_solver* s);

/* Device time of the hot launches of the last solve, by HIP events on the solver's stream (no
 * counterpart in the reference: measurement only). Profiling adds 2 events per launch. */
typedef struct fem_profile {
    uint64_t smoother_launches;  /* finest-level Chebyshev steps (operator fused with the update) */
    double smoother_seconds;
    uint64_t smoother_bytes;     /* algorithmic bytes of one step: x, b, D^-1, d in; d, x out; coefficients */
    uint64_t outer_launches;     /* FP64 operator applies of the outer PCG (q = A p) */
    double outer_seconds;
    uint64_t outer_bytes;        /* p in, q out, FP64 coefficients */
    uint64_t dofs;               /* unknowns of the finest grid, fixed ones included */
} fem_profile;
void fem_solver_set_profile(fem_solver* s, int on);
int fem_solver_profile(fem_solver* s, fem_profile* out);

/* The mean device time of the same two launches, each repeated `reps` times back to back on the
 * solver's stream between one pair of HIP events (no launch gaps inside): the finest-level
 * smoother step and the outer q = A p. Measurement only; later solves are unaffected. */
int fem_solver_time_kernels(fem_solver* s, int reps, fem_profile* out);

/* ---- the serving loop (no counterpart in the reference) ----
 * A persistent request loop over solves of uneven sizes: submit queues a solve (model, options,
 * right side), which is admitted in arrival order while the admitted solves' estimated device
 * footprint stays within hbm_budget and fewer than max_active are on the device. Setup threads
 * build each admitted solve (model, hierarchy, coarse factor) on its own stream, and the loop
 * thread advances every running solve by one PCG iteration at a time without waiting (a replayed
 * hipGraph per iteration on the solve's own stream), so the work of all running solves overlaps
 * on the device; a finished solve frees its slot and memory for the next. Each solve runs the
 * same launches as fem_solver_solve on its own stream and buffers, so its result is bitwise the
 * same alone and in any mix. Reference analogues: Solver reuse across right sides
 * (heat.rs:697-735), the cancel closure (multigrid.rs:735), refusal by size before allocation
 * (heat.rs:1180-1186). */
typedef struct fem_server fem_server;

typedef struct fem_server_options {
    uint64_t hbm_budget; /* device bytes the admitted solves may hold together (estimated footprints) */
    int32_t max_active;  /* solves admitted at once */
    int32_t device;      /* HIP device ordinal of every solve */
    int32_t setup_threads; /* host threads building admitted solves (model, hierarchy, factor); 0: 1 */
    int32_t unbatched;   /* 0 (default): continuous batching, the running solves' PCG iterations go out
                            together, each launch of a level covering that level of every solve
                            (fem_server_batching); 1: each solve launches its own iterations on its
                            own stream */
} fem_server_options;

typedef struct fem_request_info {
    double submit_time;   /* seconds since the server started */
    double admit_time;    /* admitted (slot and memory reserved) */
    double start_time;    /* set up (hierarchy built), first PCG launch */
    double finish_time;   /* solution on the host */
    uint64_t footprint;   /* estimated device bytes reserved at admission */
    int64_t device_bytes; /* measured peak device bytes of the solve's construction */
} fem_request_info;

int fem_server_new(const fem_server_options* opt, fem_server** out, char* err, size_t errlen);

/* Queues a solve. `b` holds the free unknowns' right side (copied), or NULL for the model's own
 * (fem_solver_rhs); `opt` NULL takes the defaults. Refuses a solve whose estimated footprint
 * exceeds the budget. */
int fem_server_submit(fem_server* srv, const fem_box* box, const fem_options* opt, const double* b, uint64_t* id,
                      char* err, size_t errlen);

/* Waits for solve `id` (timeout_s < 0: no limit) and takes its result: x (n free unknowns, as
 * fem_model_free_count), stats and timings. Returns 0, 1 when the solve failed (its error in
 * err), 2 on timeout (the request stays). */
int fem_server_wait(fem_server* srv, uint64_t id, double timeout_s, double* x, uint64_t n, fem_stats* st,
                    fem_request_info* info, char* err, size_t errlen);

/* Asks a queued or running solve to stop; it ends with the error "cancelled". */
int fem_server_cancel(fem_server* srv, uint64_t id);

/* The server's device memory: every admitted solve takes its memory (vectors, hierarchy, coarse
 * factor and its temporaries) from one arena of hbm_budget bytes that fem_server_new reserves, so
 * that no device allocation or free (each of which waits for the whole device) runs while solves
 * run. Gives the arena's size, the most of it in use at once, and the allocations that found no
 * room in it and were made apart (0 when the footprint estimates hold). */
int fem_server_memory(fem_server* srv, uint64_t* arena_bytes, uint64_t* peak_bytes, uint64_t* misses);

/* Continuous batching (no counterpart in the reference). Every running solve whose next PCG
 * iteration is due joins the next batch: the launches of one iteration of each member (recorded
 * once per solve at its setup) are merged in order, so that one launch runs the same kernel for
 * every member at once (the same level of each solve's V-cycle, the same PCG step), each member's
 * blocks doing exactly the work of its own launch. A batch with its readbacks is one captured
 * hipGraph, rebuilt only when the set of members changes. Solves whose iteration holds a launch
 * that cannot be batched (the hex8 marching kernels) run their own iterations as before. */
typedef struct fem_server_batch_stats {
    uint64_t batches;       /* batched PCG iterations launched */
    uint64_t iterations;    /* solve iterations they carried (members summed over batches) */
    uint64_t launches;      /* kernel launches of those batches */
    uint64_t solo_launches; /* the launches the same iterations take one solve at a time */
    uint64_t graphs;        /* batch graphs captured: a set of members due twice in a row is captured
                               once and replayed; a new set is launched directly */
    double graph_seconds;   /* loop-thread time capturing and instantiating batch graphs */
    double launch_seconds;  /* loop-thread time planning and launching batches (graph time included) */
    double finish_seconds;  /* loop-thread time finishing solves: the solution's copy, the solver's release */
} fem_server_batch_stats;
int fem_server_batching(fem_server* srv, fem_server_batch_stats* out);

/* Stops the loop (after the running solves' current step) and frees every solve. A solve that had
 * not finished ends with the error "the server stopped before the solve finished"; threads inside
 * fem_server_wait return with it before this call returns. */
void fem_server_free(fem_server* srv);

/* ---- the model alone, on the host (no device is touched) ---- */

typedef struct fem_model fem_model;

/* The numbering, free unknowns and loads of a box model, as the reference's callers build them
 * before Solver::new (heat.rs:755-877 and 503; elastic_tet.rs:88-291). */
int fem_model_new_box(const fem_box* box, fem_model** out, char* err, size_t errlen);
uint64_t fem_model_point_count(const fem_model* m);
uint64_t fem_model_unknown_count(const fem_model* m);
uint64_t fem_model_free_count(const fem_model* m);
int fem_model_free_list(const fem_model* m, int64_t* out);
int fem_model_point_grid(const fem_model* m, uint32_t* out);
/* The load over all unknowns (reference order): heat sources, body forces, face tractions. */
int fem_model_load(const fem_model* m, double* out);
/* The cell (x fastest) and the Kuhn type (0..5: chain axis orders xyz, xzy, yxz, yzx, zxy, zyx)
 * of each tetrahedron, in the reference's numbering of tetrahedra (the extrusion of
 * crates/fem/src/prism.rs:244-284 cut by split_prism, prism.rs:144-154), so that a caller can
 * give materials by position. 6 per cell; returns 1 for a hex8 model. */
int fem_model_tet_cells(const fem_model* m, uint32_t* cell, uint8_t* type);
void fem_model_free(fem_model* m);

#ifdef __cplusplus
}
#endif

#endif /* FEM_SOLVER_H */
