"""Box-grid elasticity on hex cells through the library's own box model (fem_solver_new_box with
FEM_ELASTIC_HEX8): the condensed incompatible-mode hex of each cell assembled on the host, the
reduced matrix with its grid level and cells given to the device solver, as the reference's
box-grid caller does (elastic.rs:287-547, hex.rs:199-279, multigrid.rs:458-671).

Tolerances: the reference's printed cantilever tip deflections (validation/cases/grid-cantilever,
derivation.md) to half a unit of their last printed digit (0.5e-6 mm), solved to 1e-12; random
materials with a body force and tractions against the oracle's direct solve at the 1e-8 relative
L2 parity bar; the right side and the matrix of the solver equal the host assembly bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_X = 1e-8


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# grid-cantilever: a steel beam 100 x 10 x 10 mm (E = 210 GPa, nu = 0.3) clamped at x = 0, with a
# traction of -1 MPa in z on its tip face; tip.uz is the z displacement of the tip face's centre node.
@pytest.mark.parametrize("max_cell,unknowns,printed", [(5.0, 540, -0.189395), (2.5, 3000, -0.190113),
                                                       (1.25, 19440, -0.190444), (0.625, 138720, -0.190579)])
def test_hex_cantilever_box_reproduces_the_references_tip_deflection(fem, max_cell, unknowns, printed):
    length = [100.0, 10.0, 10.0]
    n = [round(length[d] / max_cell) for d in range(3)]
    box = fem.Box("elastic_hex8", n, length, 210e3, 0.3, fixed={"x0": True}, traction={"x1": (0.0, 0.0, -1.0)})
    s = fem.Solver(box, precision="fp32", tol=1e-12)
    assert s.n == unknowns
    # the reference's routing (elastic.rs:537-547): Crossover::Flops, cubes are not flat
    entries, flops = s.crossover_counts()
    direct = entries <= 250_000_000 and flops <= 1.5e5 * unknowns
    assert s.method() == ("direct" if direct else "multigrid")
    x, st = s.solve(s.rhs())
    pg = fem.Model(box).point_grid()
    free = s.free_list()
    tip = 3 * int(np.flatnonzero((pg[:, 0] == n[0]) & (pg[:, 1] == n[1] // 2) & (pg[:, 2] == n[2] // 2))[0]) + 2
    uz = float(x[np.searchsorted(free, tip)])
    print(max_cell, s.method(), uz, st)
    assert free[np.searchsorted(free, tip)] == tip
    assert abs(uz - printed) <= 0.5e-6


@pytest.mark.parametrize("method,precision", [("direct", "fp64"), ("multigrid", "fp64"), ("multigrid", "fp32")])
def test_hex_box_matches_the_oracle(fem, oracle, method, precision):
    n, length = [24, 10, 8], [4.0, 2.0, 1.5]
    rng = np.random.default_rng(11)
    nc = int(np.prod(n))
    e = np.exp(rng.standard_normal(nc))
    nu = rng.uniform(0.15, 0.4, nc)
    body = (0.0, 0.5, -1.0)
    traction = {"x1": (0.0, 0.0, -0.25), "y1": (0.1, 0.0, 0.0)}
    box = fem.Box("elastic_hex8", n, length, e, nu, fixed={"x0": True}, body=body, traction=traction)
    (ptr, idx, val), rhs, _ = fem.hex_box_system(box)
    s = fem.Solver(box, method=method, precision=precision, tol=1e-12)
    assert np.array_equal(s.rhs(), rhs)
    sp, si, sv = s.matrix()
    assert np.array_equal(sp, ptr) and np.array_equal(si, idx) and np.array_equal(sv, val)
    pr = oracle.Problem.elastic_hex(n, length, e, nu, clamped=["x0"], body=body)
    assert np.array_equal(s.free_list(), pr.free())
    xd, _, _ = oracle.Solver(pr, "direct").solve(rhs)
    x, st = s.solve(rhs)
    print(method, precision, st)
    if method == "multigrid":
        assert st.residual <= 1e-12 and st.levels >= 2
    assert rel(x, xd) < TOL_X


def test_flat_hex_box_is_routed_direct_as_the_reference_routes_it(fem, oracle):
    n, length = [12, 12, 2], [12.0, 12.0, 0.5]  # cells 1 x 1 x 0.25: aspect 4 > MULTIGRID_ASPECT
    rng = np.random.default_rng(12)
    nc = int(np.prod(n))
    e = rng.uniform(1.0, 2.0, nc)
    box = fem.Box("elastic_hex8", n, length, e, 0.3, fixed={"x0": True}, body=(0.0, 0.0, -1.0))
    assert fem.box_aspect(box) == 4.0
    s = fem.Solver(box)
    assert s.method() == "direct"
    pr = oracle.Problem.elastic_hex(n, length, e, 0.3, clamped=["x0"], body=(0.0, 0.0, -1.0))
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    x, _ = s.solve(pr.rhs())
    assert rel(x, xd) < TOL_X
