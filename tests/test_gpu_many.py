"""Several right sides against one setup (fem_solver_solve_many): the reference's callers solve
many right sides with one Solver in turn (transient steps, heat.rs:697-735; DC terminals,
dc.rs:262-271; shift-invert blocks, modes.rs:412). Here the k solves run together, their PCG
iterations batched, with k - 1 shadow solvers sharing the hierarchy.

Tolerance: each right side's solution must equal fem_solver_solve of that right side alone bit for
bit (each right side's blocks run exactly the launches of its own solve on its own vectors), and
the model's own right side is checked against the oracle's direct solve at the 1e-8 parity bar.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = dict(method="multigrid", precision="fp32", tol=1e-10, levels=3)


def boxes(fem):
    rng = np.random.default_rng(31)
    n = [8, 4, 4]
    nt = 6 * int(np.prod(n))
    heat = fem.Box("heat_tet4", n, [2.0, 1.0, 1.0], rng.uniform(0.5, 2.0, nt), fixed={"x0": 1.0, "x1": 0.0})
    elas = fem.Box("elastic_tet10", n, [2.0, 1.0, 1.0], rng.uniform(1.0, 2.0, nt), rng.uniform(0.2, 0.35, nt),
                   fixed={"x0": True}, body=(0.0, 0.0, -1.0))
    k = np.exp(rng.standard_normal(12 * 12 * 12))
    hexb = fem.Box("heat_hex8", [12, 12, 12], [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0})
    return {"heat_tet4": heat, "elastic_tet10": elas, "heat_hex8": hexb}


def right_sides(s, k, seed):
    rng = np.random.default_rng(seed)
    b = s.rhs()
    B = [b, rng.standard_normal(s.n), 2.0 * b, np.zeros(s.n), b + 0.1 * rng.standard_normal(s.n)]
    return np.array(B[:k])


@pytest.mark.parametrize("kind", ["heat_tet4", "elastic_tet10", "heat_hex8"])
def test_many_right_sides_equal_their_solves_alone(fem, oracle, kind):
    box = boxes(fem)[kind]
    s = fem.Solver(box, **OPTS)
    B = right_sides(s, 5, 7)
    X0 = np.zeros_like(B)
    X0[4] = np.random.default_rng(8).standard_normal(s.n)  # a start value
    alone = []
    for i in range(len(B)):
        x, st = s.solve(B[i], x=X0[i])
        alone.append((x, st))
    X, stats = s.solve_many(B, X0)
    for i, (x, st) in enumerate(alone):
        assert np.array_equal(X[i], x), (kind, i)
        assert stats[i].iterations == st.iterations and stats[i].residual == st.residual
    assert stats[3].iterations == 0 and not X[3].any()  # the zero right side
    # the solver alone afterwards is unchanged, and fewer right sides reuse the shadows
    x, _ = s.solve(B[1])
    assert np.array_equal(x, alone[1][0])
    X2, _ = s.solve_many(B[:2])
    assert np.array_equal(X2[0], alone[0][0]) and np.array_equal(X2[1], alone[1][0])
    # the model's own right side against the oracle's direct solve
    length = [2.0, 1.0, 1.0]
    if kind == "heat_tet4":
        pr = oracle.Problem.heat([8, 4, 4], length, box.coef0, fixed={"x0": 1.0, "x1": 0.0}, tet4=True)
    elif kind == "elastic_tet10":
        pr = oracle.Problem.elastic(10, [8, 4, 4], length, box.coef0, box.coef1, clamped=["x0"], body=(0.0, 0.0, -1.0))
    else:
        pr = oracle.Problem.heat([12, 12, 12], [1.0, 1.0, 1.0], box.coef0, fixed={"x0": 0.0, "x1": 1.0})
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert np.linalg.norm(X[0] - xd) <= 1e-8 * np.linalg.norm(xd)


def test_many_right_sides_of_a_direct_solver(fem):
    box = boxes(fem)["heat_tet4"]
    s = fem.Solver(box, method="direct")
    B = right_sides(s, 3, 9)
    X, stats = s.solve_many(B)
    for i in range(3):
        x, _ = s.solve(B[i])
        assert np.array_equal(X[i], x)


def test_every_right_side_ends_and_the_first_error_is_raised(fem):
    """With an iteration limit too small for any of them, every right side runs to the limit and
    the error is the one fem_solver_solve gives for the first."""
    box = boxes(fem)["elastic_tet10"]
    s = fem.Solver(box, **OPTS)
    B = right_sides(s, 3, 10)[[0, 1, 2]]
    s.with_max_iter(2)
    with pytest.raises(fem.FemError) as alone:
        s.solve(B[0])
    with pytest.raises(fem.FemError) as many:
        s.solve_many(B)
    assert str(many.value) == str(alone.value)
    assert str(many.value).startswith("conjugate gradients did not converge")
