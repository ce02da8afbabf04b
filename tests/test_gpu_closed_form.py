"""The device solve against solutions the mathematics fixes (no oracle in the loop except to
number the tetrahedra of a layered model): a layered wall (exact piecewise-linear temperature),
a uniform strain (exact linear displacement) and a hanging bar (exact quadratic displacement for
tet10). Each in FP64 and in the mixed mode (FP32 V-cycle inside FP64 PCG), to 1e-12 relative
residual; the bar is the tolerance of the exact field at the nodes."""
import numpy as np
import pytest

from cases import layered_wall, layered_wall_k, uniform_strain_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["heat_hex8", "heat_tet4"])
def test_a_layered_wall_has_the_exact_interface_temperatures(fem, oracle, kind, precision):
    n, length, _ = layered_wall()
    k = layered_wall_k(oracle, kind == "heat_tet4")
    s = fem.Solver(fem.Box(kind, n, length, k, fixed={"x0": 100.0, "x1": 0.0}), method="multigrid",
                   precision=precision, tol=1e-13)
    t, st = s.solve(s.rhs())
    q = 100.0 / 3.2
    m = fem.Model(fem.Box(kind, n, length, k, fixed={"x0": 100.0, "x1": 0.0}))
    xs = m.point_grid()[s.free_list(), 0] * (length[0] / n[0])
    exact = np.where(xs <= 0.1, 100 - q * xs, np.where(xs <= 0.4, 100 - q * 0.1 - q * (xs - 0.1) / 0.1,
                                                       q * (0.6 - xs) / 2.0))
    print(kind, precision, st)
    assert np.abs(t - exact).max() < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["elastic_tet4", "elastic_tet10"])
def test_a_uniform_strain_is_exact(fem, kind, precision):
    n, length = [8, 4, 4], [2.0, 1.0, 0.8]
    e_mod, nu, g, traction = uniform_strain_box(4 if kind == "elastic_tet4" else 10, n)
    box = fem.Box(kind, n, length, e_mod, nu, fixed={"x0": True}, traction=traction)
    s = fem.Solver(box, method="multigrid", precision=precision, tol=1e-13)
    u, st = s.solve(s.rhs())
    grid = fem.Model(box).point_grid().astype(np.float64)
    h = length[0] / n[0] / (2 if kind == "elastic_tet10" else 1)
    want = np.outer(grid[:, 0] * h, g).ravel()[s.free_list()]
    print(kind, precision, st)
    assert np.abs(u - want).max() < 1e-12 * 10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_a_hanging_bar_has_the_exact_quadratic_displacement(fem, precision):
    n, length = [16, 4, 4], [4.0, 1.0, 1.0]
    e_mod, b = 2.0, 0.7
    box = fem.Box("elastic_tet10", n, length, e_mod, 0.0, fixed={"x0": True}, body=(b, 0, 0))
    s = fem.Solver(box, method="multigrid", precision=precision, tol=1e-13)
    u, st = s.solve(s.rhs())
    x = fem.Model(box).point_grid()[:, 0] * (length[0] / n[0] / 2)
    want = np.zeros((len(x), 3))
    want[:, 0] = b * (length[0] * x - x * x / 2) / e_mod
    print(precision, st)
    assert np.abs(u - want.ravel()[s.free_list()]).max() < 1e-10 * np.abs(want).max()
