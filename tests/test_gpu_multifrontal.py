"""The multifrontal LDL^T on the device (multifrontal.hip), the direct method of the assembled
boundary above 8,000 unknowns, against the oracle's multifrontal Cholesky (cholesky.rs:44-160) on
the same matrices: the same nested-dissection order and fundamental supernodes, LDL^T here and
L L^T there, so the solutions agree to rounding.

Tolerances: 1e-10 relative L2 against the oracle's direct solve, against the device's dense LDL^T
where that holds the matrix, and the true residual below 1e-12 (FP64 direct factors whose sums
run in other orders); two solves with one factor are bitwise equal; a matrix that is not positive
definite fails with the reference's text for the same column as the oracle.
"""
import numpy as np
import pytest

from test_gpu_csr import heat, rel

pytestmark = pytest.mark.gpu


def grid_of(fem, planes, pts, comp=None):
    return fem.Grid(planes, pts, np.zeros(len(pts), np.uint8) if comp is None else comp)


@pytest.mark.parametrize("n", [[24, 20, 18], [34, 34, 34]])
def test_heat_matches_the_oracle_and_the_dense_factor(fem, oracle, n):
    length = [1.0, 0.8, 0.7]
    pr, planes, pts = heat(oracle, n, length, 41)
    ptr, idx, val = pr.reduced().arrays()
    g = grid_of(fem, planes, pts)
    s = fem.CsrSolver(ptr, idx, val, "grid", g, method="direct", direct_solver="multifrontal")
    b = pr.rhs()
    x, st = s.solve(b)
    print(n, len(b), st, flush=True)
    assert st.method == "direct" and st.true_residual < 1e-12
    xd, _, _ = oracle.Solver(pr, "direct").solve(b)
    assert rel(x, xd) < 1e-10
    x2, _ = s.solve(b)
    assert np.array_equal(x, x2)
    if len(b) <= 40000:
        xdense, _ = fem.CsrSolver(ptr, idx, val, "grid", g, method="direct", direct_solver="dense").solve(b)
        assert rel(x, xdense) < 1e-10


def test_automatic_choice_takes_the_multifrontal_factor_beyond_8000_unknowns(fem, oracle):
    pr, planes, pts = heat(oracle, [22, 22, 22], [1.0, 1.0, 1.0], 42)
    ptr, idx, val = pr.reduced().arrays()
    assert len(ptr) - 1 > 8000
    s = fem.CsrSolver(ptr, idx, val, "grid", grid_of(fem, planes, pts), method="direct")
    x, st = s.solve(pr.rhs())
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert rel(x, xd) < 1e-10


def test_flat_hex_elasticity_beyond_the_dense_limit(fem, oracle):
    """A flat plate of box-grid hex elasticity (cells 1 x 1 x 0.25, clamped on its 4 edges): 50,625
    unknowns, which the reference's crossover sends to its direct factor (elastic.rs:525-547, flat),
    beyond the dense factor's 40,000. A thin plate's matrix is ill-conditioned (bending against
    membrane stiffness), so two direct factors agree to the parity bar of 1e-8, not to 1e-10."""
    n, length = [76, 76, 2], [76.0, 76.0, 0.5]
    rng = np.random.default_rng(43)
    nc = int(np.prod(n))
    e = rng.uniform(1.0, 2.0, nc)
    clamped = ["x0", "x1", "y0", "y1"]
    box = fem.Box("elastic_hex8", n, length, e, 0.3, fixed={f: True for f in clamped}, body=(0.0, 0.0, -1.0))
    s = fem.Solver(box)
    assert s.n == 50625 and s.method() == "direct"
    x, st = s.solve(s.rhs())
    print(s.n, st, flush=True)
    pr = oracle.Problem.elastic_hex(n, length, e, 0.3, clamped=clamped, body=(0.0, 0.0, -1.0))
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    # the oracle's own residual on the same matrix: a backward-stable factor leaves about
    # eps |A| |x| / |b|, which the plate's conditioning makes large; the device's must be as small
    ptr, idx, val = pr.reduced().arrays()
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr.astype(np.int64)))
    b = pr.rhs()
    ax = np.bincount(rows, weights=val * xd[idx], minlength=len(b))
    res_oracle = float(np.linalg.norm(b - ax) / np.linalg.norm(b))
    print("relative difference from the oracle", rel(x, xd), "oracle residual", res_oracle, flush=True)
    assert st.true_residual < 10 * res_oracle + 1e-14
    assert rel(x, xd) < 1e-8


def test_a_singular_matrix_fails_as_the_reference_fails(fem, oracle):
    """Heat with no fixed temperature: constants are in the null space, and the factor meets a zero
    pivot. The text and the column are the oracle's (cholesky.rs:92-99)."""
    n, length = [20, 20, 24], [1.0, 1.0, 1.0]
    k = np.exp(np.random.default_rng(44).standard_normal(int(np.prod(n))))
    pr = oracle.Problem.heat(n, length, k, fixed={})
    ptr, idx, val = pr.reduced().arrays()
    planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = pr.ijk()[pr.free()]
    with pytest.raises(oracle.OracleError) as want:
        oracle.Solver(pr, "direct")
    with pytest.raises(fem.FemError) as got:
        fem.CsrSolver(ptr, idx, val, "grid", grid_of(fem, planes, pts), method="direct", direct_solver="multifrontal")
    print(str(want.value), "|", str(got.value))
    assert str(got.value) == str(want.value)


def test_a_box_model_forced_direct_beyond_the_dense_limit(fem, oracle):
    """The box boundary's direct method (fem_solver_new_box with FEM_METHOD_DIRECT) above 8,000
    unknowns factors the model's reduced matrix (its pattern, the values probed from the FP64
    operator) by the multifrontal LDL^T: 47,915 heat unknowns, beyond the dense 40,000."""
    n, length = [36, 36, 36], [1.0, 1.0, 1.0]
    k = np.exp(np.random.default_rng(45).standard_normal(int(np.prod(n))))
    box = fem.Box("heat_hex8", n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    s = fem.Solver(box, method="direct")
    assert s.n == 47915 and s.method() == "direct"
    x, st = s.solve(s.rhs())
    print(st, flush=True)
    assert st.true_residual < 1e-12
    pr = oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert rel(x, xd) < 1e-10
