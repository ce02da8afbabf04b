"""Two of the reference's own validation cases, checked against the values its solver printed.

The derivation notes of validation/cases/grid-heated-bar and validation/cases/grid-cantilever
tabulate what the reference's grid solver gave on each mesh of a mesh study (to 7 and 6 significant
digits). The same meshes, elements, numbering and loads through this library's device solve must
reproduce those values to the printed digits: this pins the hex8 conduction operator and the box
heat model (heat.rs:1165-1278) and the condensed incompatible-mode hex with the box-grid
elasticity caller's loads and its multigrid from cells (hex.rs:199-279, elastic.rs:537-547,
elastic.rs:942-963, multigrid.rs:458-671) against the reference itself, not only against the
oracle's restatement.

Tolerances: the solves run to 1e-12 relative residual; a tabulated value v printed with d decimals
is matched to half a unit of its last digit (0.5e-6 for the bar in degC, 0.5e-6 mm for the tip).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# grid-heated-bar (derivation.md, "Discretization error and tolerance"): a square bar of side
# 0.02 m and length 0.0025 m, k = 20, q = 1e7, T = 0 on its 4 long sides; quarter symmetry in y and
# z. The solved quarter is y, z in [0, 0.01] with T = 0 on y- and z- (y+ and z+ are the symmetry
# planes, insulated); the centre of the bar is the quarter's corner node (x = 0.00125, y = z = 0.01).
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("max_cell,printed", [(0.00125, 14.779861), (0.000625, 14.745623)])
def test_heated_bar_centre_temperature_is_the_references(fem, max_cell, printed, precision):
    n = [round(0.0025 / max_cell), round(0.01 / max_cell), round(0.01 / max_cell)]
    box = fem.Box("heat_hex8", n, [0.0025, 0.01, 0.01], 20.0, fixed={"y0": 0.0, "z0": 0.0}, source=1e7)
    s = fem.Solver(box, method="multigrid", precision=precision, tol=1e-12)
    t, st = s.solve(s.rhs())
    pg = fem.Model(box).point_grid()[s.free_list()]
    at = np.flatnonzero((pg[:, 0] == n[0] // 2) & (pg[:, 1] == n[1]) & (pg[:, 2] == n[2]))
    assert at.size == 1
    centre = float(t[at[0]])
    print(max_cell, precision, centre, st)
    assert abs(centre - printed) <= 0.5e-6


# grid-cantilever (derivation.md, "Mesh study"): a steel beam 100 x 10 x 10 mm (E = 210 GPa,
# nu = 0.3) of box hex cells with incompatible modes, clamped at x = 0, traction -1 MPa in z on the
# tip face (a quarter of each tip cell face's force at its 4 points, elastic.rs:942-963); `tip.uz`
# is the z displacement of the centre node of the tip face.
@pytest.mark.parametrize("max_cell,unknowns,printed", [(5.0, 540, -0.189395), (2.5, 3000, -0.190113),
                                                       (1.25, 19440, -0.190444), (0.625, 138720, -0.190579)])
def test_hex_cantilever_tip_deflection_is_the_references(fem, oracle, max_cell, unknowns, printed):
    length = [100.0, 10.0, 10.0]
    n = [round(length[d] / max_cell) for d in range(3)]
    pr = oracle.Problem.elastic_hex(n, length, 210e3, 0.3, clamped=["x0"])
    ijk = pr.ijk()
    node_of = {tuple(p): i for i, p in enumerate(ijk)}
    force = np.zeros(3 * len(ijk))
    area = (length[1] / n[1]) * (length[2] / n[2])
    for j in range(n[1]):
        for k in range(n[2]):
            for a in (0, 1):
                for b in (0, 1):
                    force[3 * node_of[(n[0], j + a, k + b)] + 2] += -1.0 * area / 4.0
    free = pr.free()
    assert free.size == unknowns
    b = pr.rhs() + force[free]
    ptr, idx, val = pr.reduced().arrays()
    planes = [np.linspace(0.0, length[d], n[d] + 1) for d in range(3)]
    grid = fem.Grid(planes, ijk[free // 3], (free % 3).astype(np.uint8))
    # the reference's choice (solve.rs:98-183): direct up to 3,000 unknowns, multigrid above
    s = fem.CsrSolver(ptr, idx, val, "grid", grid, precision="fp32", tol=1e-12, cells=pr.cell_matrices())
    assert s.method() == ("direct" if unknowns <= 3000 else "multigrid")
    x, st = s.solve(b)
    tip = 3 * node_of[(n[0], n[1] // 2, n[2] // 2)] + 2
    uz = float(x[np.searchsorted(free, tip)])
    print(max_cell, uz, st)
    assert free[np.searchsorted(free, tip)] == tip
    assert abs(uz - printed) <= 0.5e-6
