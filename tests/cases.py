"""Closed-form cases shared by the oracle tests and the device tests."""
import numpy as np


def layered_wall(n_per=(4, 12, 8)):
    """crates/fem/tests/heat.rs:66-127: 3 layers along x (0.1 of k 1, 0.3 of k 0.1, 0.2 of k 2),
    0.1 x 0.1 across, T = 100 at x = 0 and 0 at x = 0.6. Resistance 3.2 per area."""
    nx = sum(n_per)
    k_x = np.concatenate([np.full(n_per[0], 1.0), np.full(n_per[1], 0.1), np.full(n_per[2], 2.0)])
    n = [nx, 2, 2]
    return n, [0.6, 0.1, 0.1], k_x


def layered_wall_k(oracle, tet4):
    """The conductivity of each cell (hex8, x fastest) or of each tetrahedron (by its centroid:
    the reference numbers tetrahedra by extrusion, not cell by cell)."""
    n, length, k_x = layered_wall()
    if not tet4:
        return np.tile(k_x, n[1] * n[2])
    probe = oracle.Problem.heat(n, length, np.ones(6 * int(np.prod(n))), fixed={"x0": 0.0}, tet4=True)
    cx = probe.points()[probe.cells()].mean(axis=1)[:, 0]
    return k_x[np.minimum((cx / (length[0] / n[0])).astype(int), n[0] - 1)]


def layered_wall_problem(oracle, tet4):
    n, length, _ = layered_wall()
    return oracle.Problem.heat(n, length, layered_wall_k(oracle, tet4), fixed={"x0": 100.0, "x1": 0.0}, tet4=tet4)


def uniform_strain_box(order, n):
    """A uniform strain field that vanishes on x = 0, u = (g0 x, g1 x, g2 x), on a box clamped at
    x = 0 with the tractions sigma n of its uniform stress on the other 5 faces: the discrete
    solution is that field exactly (a linear field lies in the tet4 and tet10 spaces, and the
    face loads of a uniform traction are exact)."""
    e_mod, nu = 3.0, 0.3
    g = np.array([2e-3, -1e-3, 5e-4])
    eps = np.zeros((3, 3))
    eps[0, 0] = g[0]
    eps[0, 1] = eps[1, 0] = g[1] / 2
    eps[0, 2] = eps[2, 0] = g[2] / 2
    lam = e_mod * nu / ((1 + nu) * (1 - 2 * nu))
    mu = e_mod / (2 * (1 + nu))
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    traction = {"x1": sig @ [1, 0, 0], "y0": sig @ [0, -1, 0], "y1": sig @ [0, 1, 0], "z0": sig @ [0, 0, -1],
                "z1": sig @ [0, 0, 1]}
    return e_mod, nu, g, traction
