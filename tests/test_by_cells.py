"""Box-grid hex elasticity with incompatible modes: the condensed element (hex.rs:199-279) and the
coarse matrices from cells (multigrid.rs:458-671 by_cells, modes, condense), in the oracle and
in the library's host setup. Pins: the reference's own hex.rs tests (symmetry and rigid motions,
the uniform-strain patch, one cell bending with the exact curvature), closed forms of by_cells
(rigid motions in its null space, the exact energy of a uniform strain), and the library's coarse
matrices against the oracle's bit for bit. CPU only."""
import numpy as np
import pytest

BITS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
SIZE = np.array([2.0, 0.5, 0.1])  # hex.rs:404


def nodal(f, size=SIZE):
    return np.concatenate([f(b * size) for b in BITS])


def lame(e, nu):
    return e * nu / ((1 + nu) * (1 - 2 * nu)), e / (2 * (1 + nu))


def test_the_element_is_symmetric_and_a_rigid_motion_has_no_force(oracle):
    """hex.rs:444-472 with steel (E 200, nu 0.3)."""
    k = oracle.hex_stiffness(SIZE, 200.0, 0.3)
    assert np.abs(k - k.T).max() <= 1e-12 * np.abs(k).max()
    scale = np.abs(k).max() * SIZE.max()
    for m in range(6):
        if m < 3:
            u = nodal(lambda x, m=m: np.eye(3)[m])
        else:
            w = np.eye(3)[m - 3]
            u = nodal(lambda x, w=w: np.cross(w, x))
        assert np.abs(k @ u).max() < 1e-12 * scale


def test_a_uniform_strain_has_the_exact_energy(oracle):
    """The patch test of one cell (hex.rs:474-503): a linear field's energy is eps:D:eps V."""
    e, nu = 200.0, 0.3
    k = oracle.hex_stiffness(SIZE, e, nu)
    eps = np.array([[1e-3, 2e-4, -1e-4], [2e-4, -5e-4, 3e-4], [-1e-4, 3e-4, 2e-3]])
    u = nodal(lambda x: eps @ x)
    lam, mu = lame(e, nu)
    exact = (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps)) * np.prod(SIZE)
    assert abs(u @ k @ u - exact) < 1e-12 * exact


def test_one_cell_bends_with_the_exact_curvature(oracle):
    """hex.rs:505-530: E 1000, nu 0; u_x = -kappa x (z - c/2), u_z = kappa x^2 / 2; u^T K u equals
    E I kappa^2 L."""
    e, kappa = 1000.0, 1e-3
    k = oracle.hex_stiffness(SIZE, e, 0.0)
    u = nodal(lambda x: np.array([-kappa * x[0] * (x[2] - SIZE[2] / 2), 0.0, kappa * x[0] ** 2 / 2]))
    i = SIZE[1] * SIZE[2] ** 3 / 12
    exact = e * i * kappa ** 2 * SIZE[0]
    assert abs(u @ k @ u - exact) < 1e-9 * exact


def _coarse_xyz(cp, n, length):
    """The coordinates of coarse points: the kept planes of coarsen (every other plane plus the
    last, multigrid.rs:335-345)."""
    planes = [np.linspace(0.0, length[d], n[d] + 1) for d in range(3)]
    keep = [list(range(0, n[d] + 1, 2)) + ([n[d]] if n[d] % 2 else []) for d in range(3)]
    cpl = [planes[d][keep[d]] for d in range(3)]
    return np.stack([cpl[d][cp[:, d]] for d in range(3)], axis=1)


@pytest.mark.parametrize("n,length", [([8, 6, 4], [2.0, 1.5, 1.0]), ([7, 5, 3], [1.0, 1.0, 0.3])])
def test_coarse_matrices_from_cells_hold_rigid_motions_and_uniform_strain(oracle, n, length):
    """A floating box (no support): every rigid motion of the coarse corners has no force, and a
    uniform strain's coarse corner displacements have the energy eps:D:eps V of the box (the
    trilinear fields hold a linear field exactly and its incompatible modes are 0). Odd counts
    give coarse cells of one fine layer."""
    e, nu = 1.0, 0.3
    pr = oracle.Problem.elastic_hex(n, length, e, nu)
    a, cp, cc = oracle.by_cells_step(pr)
    K = a.to_scipy()
    assert abs(K - K.T).max() <= 1e-14 * abs(K).max()
    xc = _coarse_xyz(cp, n, length)
    rows = np.arange(len(cp))
    for m in range(6):
        if m < 3:
            u = np.zeros((len(cp), 3))
            u[:, m] = 1
        else:
            u = np.cross(np.eye(3)[m - 3], xc)
        assert np.abs(K @ u[rows, cc]).max() < 1e-13 * abs(K).max()
    eps = np.array([[1e-3, 2e-4, 0.0], [2e-4, -5e-4, 3e-4], [0.0, 3e-4, 2e-3]])
    uc = (xc @ eps.T)[rows, cc]
    lam, mu = lame(e, nu)
    exact = (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps)) * np.prod(length)
    assert abs(uc @ (K @ uc) - exact) < 1e-12 * exact


def test_multigrid_from_cells_solves_a_clamped_box(oracle):
    """Solver::new with cells (multigrid.rs:84-139): two levels on 24 x 12 x 12 cells of random
    materials, the solution equal to the oracle's direct solve."""
    n, length = [24, 12, 12], [2.0, 1.0, 1.0]
    rng = np.random.default_rng(3)
    nc = int(np.prod(n))
    pr = oracle.Problem.elastic_hex(n, length, rng.uniform(1, 3, nc), rng.uniform(0.2, 0.35, nc), clamped=["x0"],
                                    body=(0, 0, -1))
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    s = oracle.Solver.with_cells(pr, "multigrid")
    assert s.levels() == 2
    x, it, res = s.solve(pr.rhs(), tol=1e-12)
    assert it <= 20
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)


def _hex_problem(oracle, n, length, seed, clamped=("x0",)):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(n))
    return oracle.Problem.elastic_hex(n, length, rng.uniform(1, 3, nc), rng.uniform(0.2, 0.35, nc),
                                      clamped=list(clamped), body=(0, 0, -1))


def _grid(fem, pr, n, length):
    planes = [np.linspace(0.0, length[d], n[d] + 1) for d in range(3)]
    free = pr.free()
    return fem.Grid(planes, pr.ijk()[free // 3], (free % 3).astype(np.uint8))


@pytest.mark.parametrize("n,length", [([24, 12, 12], [2.0, 1.0, 1.0]), ([29, 9, 11], [3.0, 0.9, 1.1])])
def test_the_library_builds_the_oracle_levels_from_cells_bit_for_bit(fem, oracle, n, length):
    """The host setup of fem_solver_new_csr with cells against the oracle's Multigrid::new with
    cells: the same prolongations and coarse matrices (pattern and FP64 values) on every level;
    odd counts give coarse cells of one fine layer and planes kept at the ends."""
    pr = _hex_problem(oracle, n, length, 4)
    a = pr.reduced()
    ptr, idx, val = a.arrays()
    ours = fem.csr_levels(ptr.astype(np.uint64), idx, val, _grid(fem, pr, n, length), cells=pr.cell_matrices())
    ref = oracle.Solver.with_cells(pr, "multigrid")
    assert ref.levels() == len(ours) >= 2
    for l, lev in enumerate(ours):
        rp, ri, rv = ref.level_matrix(l).arrays()
        assert np.array_equal(rp.astype(np.uint64), lev["a"][0])
        assert np.array_equal(ri, lev["a"][1])
        assert np.array_equal(rv, lev["a"][2])
        if l + 1 < len(ours):
            pp, pi, pv = ref.level_p(l).arrays()
            assert np.array_equal(pp.astype(np.uint64), lev["p"][0])
            assert np.array_equal(pi, lev["p"][1])
            assert np.array_equal(pv, lev["p"][2])
    # the coarse matrices differ from the Galerkin products of the same prolongations
    gal = fem.csr_levels(ptr.astype(np.uint64), idx, val, _grid(fem, pr, n, length))
    assert not np.array_equal(gal[1]["a"][2], ours[1]["a"][2])
