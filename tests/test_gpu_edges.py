"""Edge cases of the solve, as the reference's pcg defines them (multigrid.rs:692-753): a zero right
side gives x = 0 after 0 iterations (706-710); the residual test comes before the iteration limit
(722-733), so a start value that already meets the tolerance takes 0 iterations and max_iter = 0
fails only when it does not; a start value is where PCG begins. Thin boxes of one cell across, on
both boundaries (box and assembled)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def heat_box(fem, n, seed=1):
    k = np.exp(np.random.default_rng(seed).standard_normal(int(np.prod(n))))
    return fem.Box("heat_hex8", n, [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0}), k


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_a_zero_right_side_gives_zero_after_no_iteration(fem, precision):
    box, _ = heat_box(fem, [16, 16, 16])
    s = fem.Solver(box, method="multigrid", precision=precision, levels=3)
    x, st = s.solve(np.zeros(s.n), x=np.random.default_rng(2).standard_normal(s.n))
    assert st.iterations == 0 and st.residual == 0.0
    assert not x.any()


def test_a_start_value_that_meets_the_tolerance_takes_no_iteration(fem):
    box, _ = heat_box(fem, [16, 16, 16])
    s = fem.Solver(box, method="multigrid", levels=3, tol=1e-10)
    b = s.rhs()
    x, st = s.solve(b)
    assert st.iterations > 0
    x2, st2 = s.solve(b, x=x)
    assert st2.iterations == 0 and st2.residual <= 1e-10
    assert np.array_equal(x2, x)
    # max_iter = 0: the residual test comes first, so this start value still succeeds ...
    s.with_max_iter(0)
    _, st3 = s.solve(b, x=x)
    assert st3.iterations == 0
    # ... and a start value that does not meet it fails with the reference's text
    with pytest.raises(fem.FemError, match=r"^conjugate gradients did not converge: the residual is .* after 0 iterations$"):
        s.solve(b)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_pcg_starts_from_the_given_value(fem, oracle, precision):
    n = [16, 12, 8]
    box, k = heat_box(fem, n, 3)
    pr = oracle.Problem.heat(n, [1, 1, 1], k, fixed={"x0": 0.0, "x1": 1.0})
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    s = fem.Solver(box, method="multigrid", precision=precision, levels=3, tol=1e-12)
    x0 = 10.0 * np.random.default_rng(4).standard_normal(s.n)
    x, st = s.solve(pr.rhs(), x=x0)
    assert st.iterations > 0
    assert rel(x, xd) < 1e-8


@pytest.mark.parametrize("kind,n", [("heat_hex8", [24, 1, 1]), ("heat_tet4", [12, 1, 2]), ("elastic_tet10", [10, 1, 1]),
                                    ("elastic_tet4", [20, 2, 1])])
def test_thin_boxes_of_one_cell_across_match_oracle_direct(fem, oracle, kind, n):
    rng = np.random.default_rng(5)
    length = [4.0, 0.5, 0.7]
    if kind.startswith("heat"):
        count = int(np.prod(n)) * (1 if kind == "heat_hex8" else 6)
        k = rng.uniform(0.5, 2.0, count)
        pr = oracle.Problem.heat(n, length, k, fixed={"x0": 1.0, "x1": 0.0}, source=0.3, tet4=kind == "heat_tet4")
        box = fem.Box(kind, n, length, k, fixed={"x0": 1.0, "x1": 0.0}, source=0.3)
    else:
        nt = 6 * int(np.prod(n))
        e, nu = rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt)
        order = 10 if kind == "elastic_tet10" else 4
        pr = oracle.Problem.elastic(order, n, length, e, nu, clamped=["x0"], body=(0, 0, -1))
        box = fem.Box(kind, n, length, e, nu, fixed={"x0": True}, body=(0, 0, -1))
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    for method in ("direct", "multigrid"):
        s = fem.Solver(box, method=method, tol=1e-12)
        assert np.array_equal(s.free_list(), pr.free())
        x, st = s.solve(pr.rhs())
        print(kind, n, method, st)
        assert rel(x, xd) < 1e-8


def test_the_assembled_boundary_keeps_the_same_edge_semantics(fem, oracle):
    n, length = [24, 18, 14], [1.0, 1.0, 1.0]
    k = np.exp(np.random.default_rng(6).standard_normal(int(np.prod(n))))
    pr = oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    ptr, idx, val = pr.reduced().arrays()
    planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = pr.ijk()[pr.free()]
    s = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), method="multigrid",
                      tol=1e-10)
    x, st = s.solve(np.zeros(s.n), x=np.ones(s.n))
    assert st.iterations == 0 and not x.any()
    b = pr.rhs()
    x, st = s.solve(b)
    x2, st2 = s.solve(b, x=x)
    assert st2.iterations == 0 and np.array_equal(x2, x)
    s.with_max_iter(0)
    with pytest.raises(fem.FemError, match=r"after 0 iterations$"):
        s.solve(b)


def test_the_automatic_smoother_degree_is_two_for_elasticity_and_three_then_one_for_heat(fem, oracle):
    """cheb_degree 0 (the default): an elastic box smooths as with degree 2 on every level (the same
    solution bit for bit); a heat box smooths with degree 3 on the finest level and 1 below, which
    is neither 2 nor 3 everywhere. Both solve to the oracle's direct solution."""
    n = [16, 8, 8]
    nt = 6 * int(np.prod(n))
    rng = np.random.default_rng(21)
    e, nu = rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt)
    ebox = fem.Box("elastic_tet4", n, [2.0, 1.0, 1.0], e, nu, fixed={"x0": True}, body=(0, 0, -1))
    auto = fem.Solver(ebox, method="multigrid", levels=3, tol=1e-10)
    two = fem.Solver(ebox, method="multigrid", levels=3, tol=1e-10, cheb_degree=2)
    b = auto.rhs()
    xa, sa = auto.solve(b)
    x2, s2 = two.solve(b)
    assert sa.iterations == s2.iterations and np.array_equal(xa, x2)

    box, k = heat_box(fem, [24, 24, 24], seed=22)
    pr = oracle.Problem.heat([24, 24, 24], [1, 1, 1], k, fixed={"x0": 0.0, "x1": 1.0})
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    xs = {}
    for deg in (0, 2, 3):
        s = fem.Solver(box, method="multigrid", levels=3, tol=1e-12, cheb_degree=deg)
        xs[deg], st = s.solve(pr.rhs())
        print(deg, st)
        assert rel(xs[deg], xd) < 1e-8
    assert not np.array_equal(xs[0], xs[2]) and not np.array_equal(xs[0], xs[3])


@pytest.mark.parametrize("boundary", ["box", "assembled"])
def test_a_nan_right_side_runs_to_the_limit_as_the_reference(fem, oracle, boundary):
    """A NaN in the right side makes every residual NaN. The reference's loop goes on while the
    residual is not at or below tol (multigrid.rs:721-733), so it runs to max_iter and fails with the
    residual printed as Rust prints a NaN."""
    box, k = heat_box(fem, [16, 16, 16])
    if boundary == "box":
        s = fem.Solver(box, method="multigrid", levels=3)
    else:
        pr = oracle.Problem.heat([16, 16, 16], [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0})
        ptr, idx, val = pr.reduced().arrays()
        planes = [np.arange(17, dtype=np.float64) / 16.0 for _ in range(3)]
        pts = pr.ijk()[pr.free()]
        s = fem.CsrSolver(ptr, idx, val, "grid", fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), method="multigrid")
    b = s.rhs() if boundary == "box" else pr.rhs()
    b[3] = np.nan
    s.with_max_iter(4)
    with pytest.raises(fem.FemError, match="^conjugate gradients did not converge: the residual is NaN after 4 iterations$"):
        s.solve(b)
