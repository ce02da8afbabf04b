"""The oracle pinned against the reference's own known-answer tests (restated with their values)
and against what the mathematics fixes: closed forms, invariants and independent computations
(numpy / scipy), chosen so that a dropped term, a wrong sign or index or a transposed operand
fails one of them. CPU only.
"""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cases import layered_wall, layered_wall_problem, uniform_strain_box


# ---------------------------------------------------------------- sparse.rs tests
def test_the_pattern_of_two_elements_couples_each_pair_of_their_nodes(oracle):
    """crates/fem/src/sparse.rs:300-306."""
    a = oracle.Csr.pattern(3, [[0, 1], [1, 2]])
    ptr, idx, _ = a.arrays()
    assert ptr.tolist() == [0, 2, 5, 7]
    assert idx.tolist() == [0, 1, 0, 1, 2, 1, 2]


def test_the_galerkin_product_of_a_1d_laplacian_is_a_coarse_laplacian(oracle):
    """crates/fem/src/sparse.rs:308-328."""
    a = oracle.Csr.pattern(3, [[0, 1], [1, 2]])
    for e in ([0, 1], [1, 2]):
        a.add_element(e, [1.0, -1.0, -1.0, 1.0])
    p = oracle.Rect.from_arrays(3, 2, [0, 1, 3, 4], [0, 0, 1, 1], [1.0, 0.5, 0.5, 1.0])
    _, idx, val = p.galerkin(a).arrays()
    assert idx.tolist() == [0, 1, 0, 1]
    assert val.tolist() == [0.5, -0.5, -0.5, 0.5]


def test_galerkin_equals_the_triple_product(oracle):
    """Rect::galerkin (sparse.rs:231) against P^T A P in scipy on a random pattern."""
    rng = np.random.default_rng(4)
    n, m = 30, 11
    elems = rng.integers(0, n, (40, 3))
    elems = np.array([e for e in elems if len(set(e)) == 3], np.uint32)
    a = oracle.Csr.pattern(n, elems)
    for e in elems:
        q = rng.standard_normal((3, 3))
        a.add_element(e, (q + q.T).ravel())
    ptr, idx, val = [], [], []
    ptr.append(0)
    for r in range(n):
        cols = sorted(rng.choice(m, 2, replace=False))
        idx += cols
        val += list(rng.uniform(0.1, 1.0, 2))
        ptr.append(len(idx))
    p = oracle.Rect.from_arrays(n, m, ptr, idx, val)
    P = sp.csr_matrix((val, idx, ptr), shape=(n, m)).toarray()
    want = P.T @ a.to_scipy().toarray() @ P
    got = p.galerkin(a).to_scipy().toarray()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


# ---------------------------------------------------------------- order.rs / cholesky.rs tests
def laplacian(oracle, n, shift=0.01):
    """cholesky.rs:218-242 laplacian(): the 5-point Laplacian plus a shift on the diagonal."""
    elems = []
    for j in range(n):
        for i in range(n):
            if i + 1 < n:
                elems.append([i + n * j, i + 1 + n * j])
            if j + 1 < n:
                elems.append([i + n * j, i + n * (j + 1)])
    a = oracle.Csr.pattern(n * n, elems)
    for e in elems:
        a.add_element(e, [1.0, -1.0, -1.0, 1.0])
    ptr, idx, val = a.arrays()
    for r in range(n * n):
        k = ptr[r] + int(np.where(idx[ptr[r]:ptr[r + 1]] == r)[0][0])
        val[k] += shift
    a = oracle.Csr.from_arrays(n * n, ptr, idx, val)
    xyz = np.array([[k % n, k // n, 0.0] for k in range(n * n)], np.float64)
    return a, xyz


def test_the_order_is_a_permutation_with_the_middle_plane_last(oracle):
    """crates/fem/src/order.rs:186-199."""
    n = 200
    a = oracle.Csr.pattern(n, [[i, i + 1] for i in range(n - 1)])
    xyz = np.array([[i, 0.0, 0.0] for i in range(n)])
    order = oracle.nested_dissection(a, xyz)
    assert sorted(order.tolist()) == list(range(n))
    assert order[-1] == 99


def test_the_factor_solves_a_grid_laplacian_to_the_precision_of_the_arithmetic(oracle):
    """crates/fem/src/cholesky.rs:244-266."""
    a, xyz = laplacian(oracle, 40)
    order = oracle.nested_dissection(a, xyz)
    sym = oracle.Symbolic(a, order)
    chol = oracle.Cholesky(sym, a)
    x = np.sin(np.arange(a.n) * 0.37)
    b = a.mul(x)
    err = np.abs(chol.solve(b) - x).max()
    assert err < 1e-9
    assert sym.entries() < 40_000
    # the same system through scipy's sparse LU
    assert np.abs(spla.spsolve(a.to_scipy().tocsc(), b) - chol.solve(b)).max() < 1e-10


def test_a_singular_matrix_is_an_error(oracle):
    """crates/fem/src/cholesky.rs:268-281."""
    a, xyz = laplacian(oracle, 10, shift=0.0)
    sym = oracle.Symbolic(a, oracle.nested_dissection(a, xyz))
    with pytest.raises(oracle.OracleError) as e:
        oracle.Cholesky(sym, a)
    assert "not positive definite" in str(e.value) or "singular" in str(e.value)


def test_the_symbolic_count_matches_the_dense_factor(oracle):
    """Symbolic::entries (symbolic.rs) against the nonzeros of a dense Cholesky factor of the
    permuted matrix (no cancellation on a grid Laplacian)."""
    a, xyz = laplacian(oracle, 9)
    order = oracle.nested_dissection(a, xyz)
    sym = oracle.Symbolic(a, order)
    A = a.to_scipy().toarray()[np.ix_(order, order)]
    L = np.linalg.cholesky(A)
    assert sym.entries() == int(np.count_nonzero(np.abs(L) > 1e-300))


# ---------------------------------------------------------------- elements
def test_conduction_matches_gauss_integration(oracle):
    """heat.rs:1246 conduction(): K_ij = k integral of grad N_i . grad N_j over a box cell, corners
    in the order of a VTK hexahedron, against a 2 x 2 x 2 Gauss rule in numpy (exact for
    trilinear functions)."""
    vtk = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    s = np.array([0.7, 0.3, 1.3])
    k = np.array([2.0, 0.5, 1.5])
    got = oracle.conduction(s, k)
    g = np.array([-1, 1]) / math.sqrt(3.0)
    want = np.zeros((8, 8))
    for xi in itertools.product(g, g, g):
        grads = []
        for c in range(8):
            b = vtk[c]
            f = [(1 + xi[d]) / 2 if b[d] else (1 - xi[d]) / 2 for d in range(3)]
            df = [(0.5 if b[d] else -0.5) for d in range(3)]
            grads.append([df[d] * f[(d + 1) % 3] * f[(d + 2) % 3] * 2.0 / s[d] for d in range(3)])
        grads = np.array(grads)
        want += (grads * k) @ grads.T * np.prod(s) / 8.0
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_tet_gradients_invert_the_corner_matrix(oracle):
    """heat.rs tet_gradients(): the barycentric gradients and the signed volume."""
    p = np.array([[0.1, 0.0, 0.2], [1.1, 0.2, 0.0], [0.3, 0.9, 0.1], [0.2, 0.3, 1.4]])
    g, v = oracle.tet_gradients(p)
    M = np.hstack([np.ones((4, 1)), p])
    inv = np.linalg.inv(M)  # columns: coefficients of each barycentric coordinate
    assert np.abs(g - inv[1:, :].T).max() < 1e-13
    assert abs(v - np.linalg.det(M) / 6.0) < 1e-14


def test_tet_conduction_is_v_gt_k_g(oracle):
    """heat.rs:972 tet_conduction(): V G^T diag(k) G, against numpy."""
    p = np.array([[0.1, 0.0, 0.2], [1.1, 0.2, 0.0], [0.3, 0.9, 0.1], [0.2, 0.3, 1.4]])
    g, v = oracle.tet_gradients(p)
    k = np.array([2.0, 0.5, 1.5])
    got = oracle.tet_conduction(g, v, k)
    want = v * (g * k) @ g.T
    assert np.abs(got - want).max() < 1e-14 * np.abs(want).max()
    assert np.abs(got.sum(axis=1)).max() < 1e-14 * np.abs(want).max()


def test_orthotropic_stiffness_inverts_the_compliance(oracle):
    """crates/fem/src/hex.rs tests: the_orthotropic_stiffness_inverts_the_compliance."""
    c, shear = oracle.orthotropic_stiffness([140.0, 10.0, 10.0], [0.3, 0.3, 0.45], [5.0, 5.0, 3.5])
    s = np.array([[1 / 140, -0.3 / 140, -0.3 / 140], [-0.3 / 140, 0.1, -0.045], [-0.3 / 140, -0.045, 0.1]])
    assert np.abs(c @ s - np.eye(3)).max() < 1e-12
    assert shear.tolist() == [3.5, 5.0, 5.0]  # G_yz, G_xz, G_xy (hex.rs:112)


def test_unstable_poisson_ratio_is_refused_with_the_reference_text(oracle):
    with pytest.raises(oracle.OracleError) as e:
        oracle.isotropic_stiffness(1.0, 0.5)
    assert "the Poisson ratios give no stable material" in str(e.value)


def tet10_edges():
    return [[0, 1], [1, 2], [0, 2], [0, 3], [1, 3], [2, 3]]  # tet10.rs:38


def test_the_mass_is_the_exact_integral_of_the_functions(oracle):
    """crates/fem/src/tet10.rs:300-336: each mass entry against the exact integral of rho N_i N_j
    in volume coordinates, 6 V a! b! c! d! / (a + b + c + d + 3)!."""
    p = np.array([[0.0, 0.0, 0.0], [2.0, 0.1, 0.0], [0.3, 1.5, 0.2], [0.1, 0.4, 0.9]])
    m = oracle.tet10_mass(p, 3.0)
    _, vol = oracle.tet_gradients(p)

    def terms(i):
        if i < 4:
            e2 = [2 if c == i else 0 for c in range(4)]
            e1 = [1 if c == i else 0 for c in range(4)]
            return [(2.0, e2), (-1.0, e1)]
        a, b = tet10_edges()[i - 4]
        return [(4.0, [1 if c in (a, b) else 0 for c in range(4)])]

    for i in range(10):
        for j in range(10):
            exact = 0.0
            for ci, ei in terms(i):
                for cj, ej in terms(j):
                    e = [x + y for x, y in zip(ei, ej)]
                    num = np.prod([math.factorial(q) for q in e])
                    exact += ci * cj * 6.0 * vol * num / math.factorial(sum(e) + 3)
            assert abs(m[3 * i, 3 * j] - 3.0 * exact) < 1e-12, (i, j)


def _linear_field_energy_check(stiffness, p, nodes, C):
    """u^T K u = 2 V W for a linear displacement (uniform strain), W = e^T C e / 2 (Voigt with
    engineering shears), and K (rigid motion) = 0."""
    rng = np.random.default_rng(9)
    G = rng.standard_normal((3, 3)) * 1e-3
    u = (nodes @ G.T).ravel()
    e = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0], G[0, 1] + G[1, 0]])
    _, vol = p
    assert abs(u @ stiffness @ u - vol * e @ C @ e) < 1e-10 * abs(vol * e @ C @ e)
    # rigid motions: 3 translations and 3 rotations
    for mot in ([1, 0, 0], [0, 1, 0], [0, 0, 1]):
        r = np.tile(mot, len(nodes))
        assert np.abs(stiffness @ r).max() < 1e-12 * np.abs(stiffness).max()
    for W in (np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0.0]]), np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0.0]])):
        r = (nodes @ W.T).ravel()
        assert np.abs(stiffness @ r).max() < 1e-11 * np.abs(stiffness).max()
    assert np.abs(stiffness - stiffness.T).max() < 1e-12 * np.abs(stiffness).max()


def voigt(normal, shear):
    C = np.zeros((6, 6))
    C[:3, :3] = normal
    C[3:, 3:] = np.diag(shear)
    return C


def test_tet10_stiffness_is_exact_for_a_uniform_strain(oracle):
    """Tet10::stiffness (tet10.rs:90): energy of a linear field, rigid motions, symmetry."""
    p = np.array([[0.0, 0.0, 0.0], [2.0, 0.1, 0.0], [0.3, 1.5, 0.2], [0.1, 0.4, 0.9]])
    normal, shear = oracle.orthotropic_stiffness([100.0, 50.0, 20.0], [0.3, 0.2, 0.1], [30.0, 10.0, 8.0])
    k = oracle.tet10_stiffness(p, normal, shear)
    nodes = np.vstack([p, [(p[a] + p[b]) / 2 for a, b in tet10_edges()]])
    _linear_field_energy_check(k, oracle.tet_gradients(p), nodes, voigt(normal, shear))


def test_tet10_stiffness_of_a_quadratic_field(oracle):
    """A quadratic displacement u = (x^2, 0, 0) with nu = 0: energy = E integral of (2x)^2, exact
    by a degree-2 rule; the volume-coordinate integral of x^2 closes the form."""
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.2, 0.0], [0.3, 1.2, 0.1], [0.2, 0.1, 0.8]])
    normal, shear = oracle.isotropic_stiffness(5.0, 0.0)
    k = oracle.tet10_stiffness(p, normal, shear)
    nodes = np.vstack([p, [(p[a] + p[b]) / 2 for a, b in tet10_edges()]])
    u = np.zeros(30)
    u[0::3] = nodes[:, 0] ** 2
    _, vol = oracle.tet_gradients(p)
    x = p[:, 0]
    int_x2 = vol / 10.0 * (np.sum(x**2) + np.sum(np.outer(x, x)[np.triu_indices(4, 1)]))
    assert abs(u @ k @ u - 5.0 * 4.0 * int_x2) < 1e-12 * abs(5.0 * 4.0 * int_x2)


def test_tet4_stiffness_is_exact_for_a_uniform_strain(oracle):
    p = np.array([[0.1, 0.0, 0.2], [1.1, 0.2, 0.0], [0.3, 0.9, 0.1], [0.2, 0.3, 1.4]])
    normal, shear = oracle.isotropic_stiffness(210.0, 0.3)
    k = oracle.tet4_stiffness(p, normal, shear)
    _linear_field_energy_check(k, oracle.tet_gradients(p), p, voigt(normal, shear))


def test_isotropic_stiffness_is_lame(oracle):
    e, nu = 210.0, 0.3
    normal, shear = oracle.isotropic_stiffness(e, nu)
    lam = e * nu / ((1 + nu) * (1 - 2 * nu))
    mu = e / (2 * (1 + nu))
    assert np.abs(normal - (lam * np.ones((3, 3)) + 2 * mu * np.eye(3))).max() < 1e-12 * e
    assert np.abs(shear - mu).max() < 1e-12 * e


# ---------------------------------------------------------------- prism.rs split
def test_each_prism_split_fills_the_prism_and_cuts_each_face_at_its_smallest_corner(oracle):
    """crates/fem/tests/tet.rs:88-140, over all 720 orders of the 6 global indices."""
    corners = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float64)
    quads = [[0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]]
    count = 0
    for g in itertools.permutations(range(6)):
        count += 1
        tets = oracle.split_prism(g)
        total = sum(abs(np.linalg.det(corners[t[1:]] - corners[t[0]])) / 6.0 for t in tets)
        assert abs(total - 0.5) < 1e-12, g
        for q in quads:
            mn = min(q, key=lambda k: g[k])
            opp = q[(q.index(mn) + 2) % 4]
            assert any(mn in t and opp in t for t in tets), (g, q)
    assert count == 720


# ---------------------------------------------------------------- tet10 numbering and p-MG transfer
def test_quadratic_edges_cover_each_edge_once(oracle):
    """Quadratic::make (tet10.rs:244-272) on a box mesh: one node per distinct edge, numbered after
    the corners, and each tet10 lists its corners then its EDGES midpoints."""
    pr = oracle.Problem.elastic(4, [3, 2, 2], [1, 1, 1], 1.0, 0.3, clamped=["x0"])
    pts = pr.points()
    tets = pr.cells()
    edges, t10 = oracle.quadratic(pts, tets)
    want = {tuple(sorted((int(t[a]), int(t[b])))) for t in tets for a, b in tet10_edges()}
    assert len(edges) == len(want)
    assert {tuple(sorted(e)) for e in edges.tolist()} == want
    nx, ny, nz = 3, 2, 2
    assert len(edges) == (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1) - (nx + 1) * (ny + 1) * (nz + 1)
    for t, q in zip(tets, t10):
        assert q[:4].tolist() == t.tolist()
        for m, (a, b) in enumerate(tet10_edges()):
            e = edges[q[4 + m] - len(pts)]
            assert sorted(e.tolist()) == sorted([int(t[a]), int(t[b])])


def test_the_p_prolongation_interpolates_linear_fields(oracle):
    """pmultigrid.rs prolongation: a linear field on the corners is carried exactly to every
    tet10 node (edge nodes take the mean of their 2 corners)."""
    pr10 = oracle.Problem.elastic(10, [3, 2, 2], [1.5, 1, 1], 1.0, 0.3)
    pts10 = pr10.points()
    p = pr10.prolongation()
    corners = p.cols // 3
    G = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, 0.6]])
    uc = (pts10[:corners] @ G.T).ravel()
    uf = p.mul(uc)
    free = pr10.free()
    want = (pts10 @ G.T).ravel()[free]
    assert np.abs(uf - want).max() < 1e-14


def test_the_p_galerkin_coarse_matrix_is_the_tet4_matrix(oracle):
    """P^T K10 P = K4 for the same material (nested spaces: a linear field is a quadratic one)."""
    rng = np.random.default_rng(3)
    n = [3, 2, 2]
    nt = 6 * int(np.prod(n))
    e = rng.uniform(1, 2, nt)
    nu = rng.uniform(0.2, 0.35, nt)
    pr10 = oracle.Problem.elastic(10, n, [1, 1, 1], e, nu)
    pr4 = oracle.Problem.elastic(4, n, [1, 1, 1], e, nu)
    P = pr10.prolongation()
    got = P.galerkin(pr10.reduced()).to_scipy().toarray()
    want = pr4.reduced().to_scipy().toarray()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


# ---------------------------------------------------------------- closed-form solutions
@pytest.mark.parametrize("tet4", [False, True])
@pytest.mark.parametrize("method", ["direct", "multigrid"])
def test_a_layered_wall_has_the_exact_interface_temperatures(oracle, tet4, method):
    n, length, k_x = layered_wall()
    pr = layered_wall_problem(oracle, tet4)
    if method == "multigrid" and tet4:
        pytest.skip("the reference's multigrid refuses tetrahedra (solve.rs:132-137)")
    s = oracle.Solver(pr, method)
    t, _, _ = s.solve(pr.rhs(), tol=1e-13)
    q = 100.0 / 3.2
    ijk = pr.ijk()
    free = pr.free()
    x_of = {int(u): ijk[u][0] for u in free}
    xs = np.array([x_of[int(u)] for u in free]) * 0.025
    exact = np.where(xs <= 0.1, 100 - q * xs, np.where(xs <= 0.4, 100 - q * 0.1 - q * (xs - 0.1) / 0.1,
                                                       q * (0.6 - xs) / 2.0))
    assert np.abs(t - exact).max() < 1e-8


@pytest.mark.parametrize("order", [4, 10])
def test_a_uniform_strain_is_exact(oracle, order):
    n, length = [4, 3, 2], [2.0, 1.0, 0.8]
    e_mod, nu, g, traction = uniform_strain_box(order, n)
    pr = oracle.Problem.elastic(order, n, length, e_mod, nu, clamped=["x0"], traction=traction)
    s = oracle.Solver(pr, "direct")
    u, _, _ = s.solve(pr.rhs())
    pts = pr.points()
    want = np.outer(pts[:, 0], g).ravel()[pr.free()]
    assert np.abs(u - want).max() < 1e-12


def test_a_hanging_bar_has_the_exact_quadratic_displacement(oracle):
    """nu = 0, body force b along x, clamped at x = 0, free elsewhere: u_x = b (L x - x^2 / 2) / E,
    quadratic, so tet10 is exact at every node."""
    n, length = [6, 2, 2], [3.0, 0.5, 0.5]
    e_mod, b = 2.0, 0.7
    pr = oracle.Problem.elastic(10, n, length, e_mod, 0.0, clamped=["x0"], body=(b, 0, 0))
    u, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    pts = pr.points()
    x = pts[:, 0]
    want = np.zeros((len(pts), 3))
    want[:, 0] = b * (length[0] * x - x * x / 2) / e_mod
    assert np.abs(u - want.ravel()[pr.free()]).max() < 1e-11 * np.abs(want).max()


# ---------------------------------------------------------------- solvers against scipy
@pytest.mark.parametrize("kind", ["heat_hex8", "elastic_tet10"])
def test_oracle_solvers_agree_with_scipy(oracle, kind):
    rng = np.random.default_rng(21)
    if kind == "heat_hex8":
        n = [10, 9, 8]
        pr = oracle.Problem.heat(n, [1, 1, 1], np.exp(rng.standard_normal(int(np.prod(n)))),
                                 fixed={"x0": 0.0, "x1": 1.0})
    else:
        n = [6, 3, 3]
        nt = 6 * int(np.prod(n))
        pr = oracle.Problem.elastic(10, n, [2, 1, 1], rng.uniform(1, 2, nt), 0.3, clamped=["x0"], body=(0, 0, -1))
    b = pr.rhs()
    want = spla.spsolve(pr.reduced().to_scipy().tocsc(), b)
    for method in ("direct", "multigrid"):
        x, it, res = oracle.Solver(pr, method, crossover=0).solve(b, tol=1e-12)
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-9, method
        if method == "multigrid":
            assert res < 1e-12 and it > 0


def test_the_hierarchy_is_galerkin(oracle):
    """multigrid.rs:109: each coarse matrix is P^T A P of the level above (scipy product)."""
    rng = np.random.default_rng(5)
    n = [20, 20, 20]  # 9261 points: 2 levels above the reference's coarsest size (3000)
    pr = oracle.Problem.heat(n, [1, 1, 1], np.exp(rng.standard_normal(int(np.prod(n)))), fixed={"x0": 0.0})
    s = oracle.Solver(pr, "multigrid", crossover=0)
    assert s.levels() >= 2
    for l in range(s.levels() - 1):
        A = s.level_matrix(l).to_scipy()
        ptr, idx, val = s.level_p(l).arrays()
        P = sp.csr_matrix((val, idx.astype(np.int64), ptr), shape=(s.level_n(l), s.level_n(l + 1)))
        want = (P.T @ A @ P).toarray()
        got = s.level_matrix(l + 1).to_scipy().toarray()
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_nonconvergence_has_the_reference_message(oracle):
    """multigrid.rs:730-733: "conjugate gradients did not converge: the residual is {:.2e} after
    {} iterations"."""
    rng = np.random.default_rng(1)
    n = [20, 20, 20]
    pr = oracle.Problem.heat(n, [1, 1, 1], np.exp(rng.standard_normal(8000)), fixed={"x0": 0.0, "x1": 1.0})
    with pytest.raises(oracle.OracleError) as e:
        oracle.Solver(pr, "multigrid", crossover=0).solve(pr.rhs(), max_iter=2, tol=1e-14)
    msg = str(e.value)
    assert msg.startswith("conjugate gradients did not converge: the residual is ")
    assert msg.endswith(" after 2 iterations")
    mant, ex = msg.split("residual is ")[1].split(" ")[0].split("e")
    assert len(mant.split(".")[1]) == 2 and not ex.startswith("+") and not ex.startswith("-0")
