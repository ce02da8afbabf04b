"""The assembled boundary on the device (fem_solver_new_csr): Solver::new(a, unknowns, ...) of
solve.rs:92-183 for the oracle's assembled matrices, each solve against the oracle's direct
(multifrontal Cholesky) solution within 1e-8 relative L2, the method chosen as the reference
chooses it, and the hierarchy's level sizes equal to the oracle's (the levels of the same
coarsen and Galerkin products)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_X = 1e-8


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def heat(oracle, n, length, seed):
    k = np.exp(np.random.default_rng(seed).standard_normal(int(np.prod(n))))
    pr = oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = pr.ijk()[pr.free()]
    return pr, planes, pts


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_grid_multigrid_on_an_assembled_heat_matrix_matches_oracle_direct(fem, oracle, precision):
    n, length = [24, 18, 14], [1.0, 0.8, 0.7]
    pr, planes, pts = heat(oracle, n, length, 5)
    a = pr.reduced()
    ptr, idx, val = a.arrays()
    comp = np.zeros(len(pts), np.uint8)
    s = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, comp), method="multigrid", precision=precision,
                      tol=1e-12)
    assert s.method() == "multigrid"
    ref = oracle.Solver.grid(a, planes, pts, comp, method="multigrid")
    for l in range(ref.levels()):
        assert s.level_info(l)[0] == ref.level_n(l)
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    x, st = s.solve(pr.rhs())
    print(precision, st)
    assert st.levels == ref.levels()
    assert st.true_residual < 1e-11
    assert rel(x, xd) < TOL_X
    # the FP64 operator is the matrix itself
    v = np.random.default_rng(1).standard_normal(s.n)
    assert rel(s.apply(v), a.mul(v)) < 1e-14


def test_small_assembled_matrices_take_the_direct_solver(fem, oracle):
    n, length = [10, 8, 6], [1.0, 1.0, 1.0]
    pr, planes, pts = heat(oracle, n, length, 2)
    ptr, idx, val = pr.reduced().arrays()
    g = fem.Grid(planes, pts, np.zeros(len(pts), np.uint8))
    s = fem.CsrSolver(ptr, idx, val, "grid", g)  # 567 unknowns <= crossover 3000 (solve.rs:158)
    assert s.method() == "direct"
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    x, st = s.solve(pr.rhs())
    assert st.iterations == 0 and st.residual < 1e-12
    assert rel(x, xd) < 1e-10


def test_points_unknowns_are_direct_only(fem, oracle):
    """A mesh of tetrahedra given as points (solve.rs:132-137): direct, and multigrid refused with
    the reference's text."""
    n, length = [10, 3, 3], [100.0, 10.0, 10.0]
    nt = 6 * int(np.prod(n))
    pr = oracle.Problem.elastic(4, n, length, np.full(nt, 210e3), np.full(nt, 0.3), clamped=["x0"],
                                traction={"z1": (0, 0, -1.0)})
    ptr, idx, val = pr.reduced().arrays()
    with pytest.raises(fem.FemError, match="^multigrid needs the box grid: a mesh of tetrahedra has the direct solver$"):
        fem.CsrSolver(ptr, idx, val, "points", method="multigrid")
    s = fem.CsrSolver(ptr, idx, val, "points")
    assert s.method() == "direct"
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    x, _ = s.solve(pr.rhs())
    assert rel(x, xd) < 1e-10


def quadratic(oracle, n, length, seed):
    nt = 6 * int(np.prod(n))
    rng = np.random.default_rng(seed)
    pr = oracle.Problem.elastic(10, n, length, rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt),
                                clamped=["x0"], body=(0, 0, -1))
    p = pr.prolongation()
    pp, pi, pv = p.arrays()
    return pr, (pp, pi, pv, p.cols)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_p_multigrid_on_an_assembled_tet10_matrix_matches_oracle_direct(fem, oracle, precision):
    """Unknowns::Quadratic: the caller's prolongation (pmultigrid.rs:104-151, made by the oracle),
    its Galerkin product as the linear level, the dense LDL^T there."""
    n, length = [8, 4, 4], [2.0, 1.0, 1.0]
    pr, p = quadratic(oracle, n, length, 3)
    ptr, idx, val = pr.reduced().arrays()
    s = fem.CsrSolver(ptr, idx, val, "quadratic", p=p, method="multigrid", precision=precision, tol=1e-12)
    assert s.method() == "multigrid"
    assert s.level_info(1)[0] == p[3]
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    x, st = s.solve(pr.rhs())
    print(precision, st)
    assert st.levels == 2
    assert rel(x, xd) < TOL_X


def test_p_multigrid_continues_with_grid_levels_below_the_linear_level(fem, oracle):
    """With the linear unknowns' grid given, the reference's grid coarsening continues below the
    linear level (here forced by a small dense limit)."""
    n, length = [16, 8, 8], [2.0, 1.0, 1.0]
    pr, p = quadratic(oracle, n, length, 4)
    ptr, idx, val = pr.reduced().arrays()
    # the linear unknowns: the free corner unknowns 3 v + c (v < corners) in free order
    corners = pr.sizes.corners
    free = pr.free()
    lin = free[free < 3 * corners]
    assert len(lin) == p[3]
    h = [length[d] / n[d] for d in range(3)]
    xyz = pr.points()[lin // 3]
    pts = np.stack([np.rint(xyz[:, d] / h[d]) for d in range(3)], axis=1).astype(np.uint32)
    planes = [h[d] * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    cg = fem.Grid(planes, pts, (lin % 3).astype(np.uint8))
    s = fem.CsrSolver(ptr, idx, val, "quadratic", p=p, coarse_grid=cg, method="multigrid", tol=1e-12,
                      coarse_max=1500)
    x, st = s.solve(pr.rhs())
    print(st)
    assert st.levels >= 3 and st.coarse_unknowns <= 1500
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert rel(x, xd) < TOL_X


def test_assembled_solves_are_bitwise_reproducible_and_report_the_reference_errors(fem, oracle):
    n, length = [24, 18, 14], [1.0, 1.0, 1.0]
    pr, planes, pts = heat(oracle, n, length, 7)
    ptr, idx, val = pr.reduced().arrays()
    g = fem.Grid(planes, pts, np.zeros(len(pts), np.uint8))
    s = fem.CsrSolver(ptr, idx, val, "grid", g, method="multigrid", tol=1e-10)
    b = pr.rhs()
    x1, s1 = s.solve(b)
    x2, s2 = s.solve(b)
    assert s1.iterations == s2.iterations and np.array_equal(x1, x2)
    s.with_max_iter(1)
    with pytest.raises(fem.FemError, match=r"^conjugate gradients did not converge: the residual is .* after 1 iterations$"):
        s.solve(b)
    s.with_max_iter(500)
    with pytest.raises(fem.FemError, match="^cancelled$"):
        s.solve(b, cancelled=lambda: True)


@pytest.mark.parametrize("kind,n", [("heat_hex8", [8, 6, 5]), ("heat_tet4", [6, 5, 4]), ("elastic_tet4", [6, 4, 3]),
                                    ("elastic_tet10", [4, 3, 3]), ("elastic_tet10", [5, 2, 3])])
def test_the_exported_matrix_is_the_reference_reduced_matrix(fem, oracle, kind, n):
    """fem_solver_matrix of a box model against the oracle's Csr::submatrix(free): the same pattern
    bit for bit, the same values to 1e-12; and the assembled boundary on that matrix solves as the
    box does."""
    rng = np.random.default_rng(8)
    length = [1.2, 1.0, 0.9]
    if kind.startswith("heat"):
        count = int(np.prod(n)) * (1 if kind == "heat_hex8" else 6)
        k = rng.uniform(0.5, 2.0, count)
        pr = oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0}, tet4=kind == "heat_tet4")
        box = fem.Box(kind, n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    else:
        nt = 6 * int(np.prod(n))
        e, nu = rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt)
        pr = oracle.Problem.elastic(10 if kind == "elastic_tet10" else 4, n, length, e, nu, clamped=["x0"],
                                    body=(0, 0, -1))
        box = fem.Box(kind, n, length, e, nu, fixed={"x0": True}, body=(0, 0, -1))
    s = fem.Solver(box, method="multigrid")
    ptr, idx, val = s.matrix()
    rp, ri, rv = pr.reduced().arrays()
    assert np.array_equal(ptr, rp.astype(np.uint64))
    assert np.array_equal(idx, ri)
    assert rel(val, rv) < 1e-12
    if kind == "heat_hex8":
        planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
        pts = fem.Model(box).point_grid()[s.free_list()]
        c = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(s.n, np.uint8)), method="direct")
        x, _ = c.solve(s.rhs())
        xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
        assert rel(x, xd) < 1e-10


def test_unknowns_numbered_far_apart_solve_as_the_grid_numbering_does(fem, oracle):
    """A random renumbering of 76,824 unknowns: every slice of 64 rows then reaches columns more than
    2^16 apart, so the device keeps 32-bit columns (the grid numbering fits 16-bit offsets from each
    slice's first column). Both numberings build the same levels up to the numbering and give the
    same solution."""
    import scipy.sparse as sp

    n, length = [48, 48, 32], [1.0, 1.0, 0.7]
    pr, planes, pts = heat(oracle, n, length, 11)
    ptr, idx, val = pr.reduced().arrays()
    m = len(ptr) - 1
    assert m > 65536
    a = sp.csr_matrix((val, idx, ptr), shape=(m, m))
    perm = np.random.default_rng(12).permutation(m)  # new unknown k is old unknown perm[k]
    ap = a[perm][:, perm].tocsr()
    ap.sort_indices()
    b = pr.rhs()
    comp = np.zeros(m, np.uint8)
    s0 = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, comp), method="multigrid", tol=1e-12)
    s1 = fem.CsrSolver(ap.indptr.astype(np.int64), ap.indices.astype(np.uint32), ap.data, "grid",
                       fem.Grid(planes, pts[perm], comp), method="multigrid", tol=1e-12)
    x0, st0 = s0.solve(b)
    x1, st1 = s1.solve(b[perm])
    print(st0, st1)
    assert abs(st0.iterations - st1.iterations) <= 1
    back = np.empty(m)
    back[perm] = x1
    assert rel(back, x0) < 1e-9
    v = np.random.default_rng(13).standard_normal(m)
    assert rel(s1.apply(v[perm]), (a @ v)[perm]) < 1e-14


def test_p_multigrid_with_a_large_linear_level_and_no_grid(fem, oracle):
    """The reference's tet10 caller (elastic_tet.rs:294-313) passes coordinates only, so the linear
    level has no grid below it and the reference factors it whole (pmultigrid.rs:46-60). Here it
    holds 12,168 unknowns, above the default coarse_max of 6,000: the device factors it dense. The
    solution against the oracle's own p-multigrid PCG (multifrontal Cholesky on the linear level) to
    1e-12."""
    n, length = [24, 12, 12], [2.0, 1.0, 1.0]
    pr, p = quadratic(oracle, n, length, 6)
    assert p[3] == 12168
    ptr, idx, val = pr.reduced().arrays()
    s = fem.CsrSolver(ptr, idx, val, "quadratic", p=p, method="multigrid", precision="fp32", tol=1e-12)
    assert s.level_info(1)[0] == p[3]
    b = pr.rhs()
    x, st = s.solve(b)
    print(st, flush=True)
    assert st.levels == 2 and st.coarse_unknowns == p[3]
    assert st.true_residual < 1e-11
    xo, it, res = oracle.Solver(pr, "multigrid").solve(b, tol=1e-12)
    print("oracle p-multigrid iterations", it, flush=True)
    assert rel(x, xo) < TOL_X


def test_a_direct_solve_of_about_34k_unknowns_matches_oracle_direct(fem, oracle):
    """Direct solves up to the reference's TRANSIENT_DIRECT_LIMIT of 40,000 unknowns (solve.rs:24):
    33,759 heat unknowns (32^3 cells) forced direct through the assembled boundary, against the
    oracle's multifrontal Cholesky."""
    n, length = [32, 32, 32], [1.0, 1.0, 1.0]
    pr, planes, pts = heat(oracle, n, length, 21)
    ptr, idx, val = pr.reduced().arrays()
    assert len(ptr) - 1 == 33759
    s = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), method="direct")
    assert s.method() == "direct"
    b = pr.rhs()
    x, st = s.solve(b)
    print(st, flush=True)
    xd, _, _ = oracle.Solver(pr, "direct").solve(b)
    assert st.iterations == 0 and st.residual < 1e-12
    assert rel(x, xd) < 1e-10


def test_a_dense_direct_solve_beyond_its_limit_is_refused(fem, oracle):
    """The dense factor holds 40,000 unknowns; asked for beyond that, it refuses with the counts
    (the automatic choice takes the multifrontal factor there, test_gpu_multifrontal.py)."""
    n, length = [34, 34, 34], [1.0, 1.0, 1.0]
    pr, planes, pts = heat(oracle, n, length, 22)
    ptr, idx, val = pr.reduced().arrays()
    with pytest.raises(fem.FemError, match="^the direct solve holds at most 40000 unknowns; this matrix has 40425$"):
        fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), method="direct",
                      direct_solver="dense")


def _oracle_counts(oracle, a, xyz):
    s = oracle.Symbolic(a, oracle.nested_dissection(a, xyz))
    return s.entries(), s.flops()


@pytest.mark.parametrize("n,flat,want", [(24, False, "direct"), (30, False, "multigrid"), (30, True, "direct")])
def test_the_flops_crossover_routes_an_assembled_grid_as_the_reference(fem, oracle, n, flat, want):
    """Crossover::Flops { flops: 1.5e5, entries: 2.5e8, flat } (elastic.rs:541 with DIRECT_FLOPS and
    DIRECT_ENTRIES): the nested-dissection factor of 24^3 heat cells costs 8.5e4 flops per unknown
    (direct), of 30^3 cells 1.7e5 (multigrid, unless the cells are flat). The counts the library
    compared equal the oracle's, the method follows solve.rs:158-170, and the solve matches."""
    length = [1.0, 1.0, 1.0]
    pr, planes, pts = heat(oracle, [n, n, n], length, 4)
    a = pr.reduced()
    ptr, idx, val = a.arrays()
    xyz = np.stack([planes[d][pts[:, d]] for d in range(3)], axis=1)
    e, f = _oracle_counts(oracle, a, xyz)
    cross = fem.Flops(1.5e5, 250_000_000, flat)
    assert (e <= cross.entries and (flat or f <= cross.flops * a.n)) == (want == "direct")
    s = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), crossover=cross,
                      tol=1e-12)
    assert s.method() == want
    assert s.crossover_counts() == (e, f)
    x, st = s.solve(pr.rhs())
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert rel(x, xd) < TOL_X


@pytest.mark.parametrize("n,cross,want", [
    ([16, 16, 1], dict(flops=float("inf"), entries=250_000_000, flat=True), "direct"),
    ([16, 16, 1], dict(flops=float("inf"), entries=1_000_000, flat=True), "multigrid"),
    ([8, 8, 8], dict(flops=1.5e5, entries=250_000_000, flat=False), "multigrid"),
])
def test_the_flops_crossover_routes_a_tet10_box_as_the_reference(fem, oracle, n, cross, want):
    """The tet10 caller's crossover (elastic_tet.rs:300-312): on flat cells Flops { flops: inf,
    entries, flat: true } (direct up to `entries` factor entries), else the flop test. A flat plate of
    16 x 16 x 1 cells 4 times thinner than wide (9,504 unknowns, 1.7e6 factor entries) is direct, or
    multigrid with a limit of 1e6 entries; a cube of 8^3 cells (2.3e5 flops per unknown) is multigrid.
    The counts equal the oracle's; each solve matches the oracle's direct solve."""
    nt = 6 * int(np.prod(n))
    length = [float(n[0]), float(n[1]), 0.25 * n[2] if n[2] == 1 else float(n[2])]
    rng = np.random.default_rng(17)
    e_, nu = rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt)
    pr = oracle.Problem.elastic(10, n, length, e_, nu, clamped=["x0"], body=(0, 0, -1))
    box = fem.Box("elastic_tet10", n, length, e_, nu, fixed={"x0": True}, body=(0, 0, -1))
    a = pr.reduced()
    e, f = _oracle_counts(oracle, a, pr.points()[pr.free() // 3])
    c = fem.Flops(**cross)
    assert (e <= c.entries and (c.flat or f <= c.flops * a.n)) == (want == "direct")
    s = fem.Solver(box, crossover=c, tol=1e-12)
    assert s.method() == want
    assert s.crossover_counts() == (e, f)
    x, st = s.solve(s.rhs())
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert rel(x, xd) < TOL_X


def test_cells_of_points_unknowns_are_refused(fem, oracle):
    pr, planes, pts = heat(oracle, [6, 5, 4], [1.0, 1.0, 1.0], 2)
    ptr, idx, val = pr.reduced().arrays()
    with pytest.raises(fem.FemError, match="^cells give the coarse matrices of grid unknowns"):
        fem.CsrSolver(ptr, idx, val, "points", cells=(np.zeros((1, 3)), np.zeros((1, 576))))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_box_grid_hex_elasticity_with_cells_matches_oracle_direct(fem, oracle, precision):
    """The reference's box-grid elasticity caller (elastic.rs:537-547) passes its cells: the grid
    hierarchy takes the by_cells coarse matrices (multigrid.rs:458-671). 24 x 12 x 12 hex cells of
    random materials, clamped at x = 0 under a body force, 12,168 unknowns."""
    n, length = [24, 12, 12], [2.0, 1.0, 1.0]
    rng = np.random.default_rng(9)
    nc = int(np.prod(n))
    pr = oracle.Problem.elastic_hex(n, length, rng.uniform(1, 3, nc), rng.uniform(0.2, 0.35, nc), clamped=["x0"],
                                    body=(0, 0, -1))
    ptr, idx, val = pr.reduced().arrays()
    planes = [np.linspace(0.0, length[d], n[d] + 1) for d in range(3)]
    free = pr.free()
    grid = fem.Grid(planes, pr.ijk()[free // 3], (free % 3).astype(np.uint8))
    s = fem.CsrSolver(ptr, idx, val, "grid", grid, method="multigrid", precision=precision, tol=1e-12,
                      cells=pr.cell_matrices())
    assert s.level_info(1)[0] == oracle.Solver.with_cells(pr, "multigrid").level_n(1)
    x, st = s.solve(pr.rhs())
    print(precision, st, flush=True)
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert st.true_residual < 1e-11
    assert rel(x, xd) < TOL_X
