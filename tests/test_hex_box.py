"""Box-grid elasticity on hex cells, host side (no device): the library's condensed
incompatible-mode hex (hex.rs:199-279) and the reduced system that fem_solver_new_box assembles
for an elastic_hex8 box (elastic.rs:287-547: Csr::pattern + add_element in cell order,
submatrix(free), the body force V/8 per corner, a quarter of each cell face's traction force on
its 4 points, elastic.rs:942-963).

Tolerance: bit for bit. The element against the oracle's restatement (itself pinned by the
reference's hex tests: rigid modes, uniform strain, one-cell bending, in test_by_cells.py); the
matrix against the oracle's assembly of the same box; the right side against the oracle's body
load with the tractions added in the reference's order (faces in load order, each face's cells in
cell order, a cell face's points in the order of grid.rs:272-289).
"""
import numpy as np
import pytest

FACE_AXIS = {"x": 0, "y": 1, "z": 2}


@pytest.mark.parametrize("size", [(1.0, 1.0, 1.0), (2.0, 0.5, 0.1), (0.3, 0.7, 1.9)])
@pytest.mark.parametrize("e,nu", [(210e3, 0.3), (1.0, 0.0), (5.0, 0.45), (1.0, -0.5)])
def test_hex_element_is_the_oracles_bit_for_bit(fem, oracle, size, e, nu):
    k = fem.hex_element(size, e, nu)
    assert np.array_equal(k, oracle.hex_stiffness(size, e, nu).reshape(24, 24))
    # symmetric to rounding: the condensation K_uu - K_um K_mm^-1 K_mu is not symmetrized
    assert np.abs(k - k.T).max() <= 1e-12 * np.abs(k).max()


def traction_load(pr, n, length, traction):
    """The oracle's body load plus the face tractions in the reference's order."""
    load = pr.load().copy()
    node = {tuple(p): i for i, p in enumerate(pr.ijk())}
    planes = [np.arange(n[d] + 1, dtype=np.float64) * (length[d] / n[d]) for d in range(3)]
    for face, t in traction:
        ax = FACE_AXIS[face[0]]
        a, b = [d for d in range(3) if d != ax]
        top = face[1] == "1"
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(n[0]):
                    c = [i, j, k]
                    if c[ax] != (n[ax] - 1 if top else 0):
                        continue
                    area = (planes[a][c[a] + 1] - planes[a][c[a]]) * (planes[b][c[b] + 1] - planes[b][c[b]])
                    for qa, qb in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = list(c)
                        p[ax] = n[ax] if top else 0
                        p[a] += qa
                        p[b] += qb
                        for d in range(3):
                            load[3 * node[tuple(p)] + d] += t[d] * area / 4.0
    return load


@pytest.mark.parametrize("n,length,clamped,traction", [
    ([5, 3, 4], [2.0, 0.7, 1.1], ["x0"], [("x1", (0.5, -0.25, 1.5)), ("z1", (0.0, 0.0, -2.0))]),
    ([4, 4, 2], [1.0, 1.0, 0.3], ["z0", "y1"], [("x0", (1.0, 0.0, 0.0)), ("x1", (-1.0, 0.5, 0.0))]),
    ([1, 1, 1], [1.0, 2.0, 3.0], ["x0"], []),
    ([7, 1, 1], [7.0, 1.0, 1.0], ["x0", "x1"], [("y1", (0.0, 1.0, 0.0))]),
])
def test_hex_box_system_is_the_oracles_bit_for_bit(fem, oracle, n, length, clamped, traction):
    rng = np.random.default_rng(3)
    nc = int(np.prod(n))
    e = rng.uniform(1.0, 3.0, nc)
    nu = rng.uniform(0.1, 0.4, nc)
    body = (0.3, -1.0, 2.0)
    box = fem.Box("elastic_hex8", n, length, e, nu, fixed={f: True for f in clamped}, body=body,
                  traction=dict(traction))
    (ptr, idx, val), rhs, aspect = fem.hex_box_system(box)
    pr = oracle.Problem.elastic_hex(n, length, e, nu, clamped=clamped, body=body)
    optr, oidx, oval = pr.reduced().arrays()
    assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx) and np.array_equal(val, oval)
    free = pr.free()
    assert np.array_equal(rhs, traction_load(pr, n, length, traction)[free])
    model = fem.Model(box)
    assert np.array_equal(model.free_list(), free)
    assert np.array_equal(model.point_grid(), pr.ijk())
    assert aspect == pytest.approx(fem.box_aspect(box), rel=0, abs=0)


def test_hex_box_materials_are_checked_with_the_references_texts(fem):
    with pytest.raises(fem.FemError, match="^the Poisson ratios give no stable material"):
        fem.hex_box_system(fem.Box("elastic_hex8", [2, 1, 1], [2.0, 1.0, 1.0], 1.0, 0.5, fixed={"x0": True}))
    with pytest.raises(fem.FemError, match="^the moduli must be above 0 and the Poisson ratios finite$"):
        fem.hex_box_system(fem.Box("elastic_hex8", [2, 1, 1], [2.0, 1.0, 1.0], 0.0, 0.3, fixed={"x0": True}))
    with pytest.raises(fem.FemError, match="^the Poisson ratios are missing$"):
        fem.hex_box_system(fem.Box("elastic_hex8", [2, 1, 1], [2.0, 1.0, 1.0], 1.0, None, fixed={"x0": True}))


def test_hex_box_aspect_is_the_references_flatness(fem):
    assert fem.box_aspect(fem.Box("elastic_hex8", [4, 4, 1], [4.0, 4.0, 0.25], 1.0, 0.3)) == 4.0
    assert fem.box_aspect(fem.Box("elastic_hex8", [3, 3, 3], [1.0, 1.0, 1.0], 1.0, 0.3)) < 1.0 + 1e-12


@pytest.mark.parametrize("size,alpha", [((2.0, 0.5, 0.1), 12e-6), ((1.0, 1.0, 1.0), 1.0), ((0.3, 0.7, 1.9), 2e-5)])
def test_hex_thermal_load_is_free_expansion(fem, size, alpha):
    """The load of a unit temperature change (hex.rs:199-279, the modes' share condensed): a free
    cell under a uniform change dt takes the strain alpha dt on each axis with no stress, so the
    corner displacements u = alpha dt x solve K u = load dt exactly; a closed form (a sign, a factor
    or a dropped condensation term fails it). The stiffness is the element's without expansion."""
    k, load = fem.hex_element(size, 3.0, 0.3, alpha)
    assert np.array_equal(k, fem.hex_element(size, 3.0, 0.3))
    bits = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    u = np.array([alpha * b[d] * size[d] for b in bits for d in range(3)])
    f = load @ np.ones(8)
    assert np.abs(k @ u - f).max() <= 1e-12 * np.abs(f).max()
    # and each corner's column is the load of a change at that corner alone: the 8 sum to the free expansion
    assert np.abs(load.sum(axis=1) - f).max() <= 1e-15 * np.abs(f).max()


VTK_BITS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _corners(size):
    return np.array([[b[d] * size[d] for d in range(3)] for b in VTK_BITS])


def test_strain_stress_of_closed_forms(fem):
    """hex.rs:313-340 on fields with known answers: a uniform strain gives that strain and D times
    it at any point (the modes stay 0); a free expansion alpha dT x and a rigid rotation give no
    stress."""
    size, e, nu, alpha = [2.0, 1.0, 0.5], 200.0, 0.3, 1e-3
    x = _corners(size)
    grad = np.array([[1e-3, 2e-3, -1e-3], [0.5e-3, -2e-3, 1e-3], [0.0, 3e-3, 1.5e-3]])
    u = (x @ grad.T).ravel()
    lam, mu = e * nu / ((1 + nu) * (1 - 2 * nu)), e / (2 * (1 + nu))
    eps = np.array([grad[0, 0], grad[1, 1], grad[2, 2], grad[1, 2] + grad[2, 1], grad[0, 2] + grad[2, 0],
                    grad[0, 1] + grad[1, 0]])
    tr = eps[:3].sum()
    sig = np.concatenate([2 * mu * eps[:3] + lam * tr, mu * eps[3:]])
    for xi in ([0.0, 0.0, 0.0], [0.3, -0.7, 1.0], [-1.0, -1.0, -1.0]):
        strain, stress = fem.hex_strain_stress(size, e, nu, alpha, xi, u, np.zeros(8))
        assert np.allclose(strain, eps, rtol=0, atol=1e-15) and np.allclose(stress, sig, rtol=1e-12, atol=1e-12)
    dt = 50.0
    free = (alpha * dt * x).ravel()
    _, stress = fem.hex_strain_stress(size, e, nu, alpha, [0.2, 0.1, -0.4], free, np.full(8, dt))
    assert np.abs(stress).max() < 1e-12
    w = np.array([0.3, -0.2, 0.1])
    rot = np.cross(w, x).ravel()
    strain, stress = fem.hex_strain_stress(size, e, nu, alpha, [0.5, 0.5, 0.5], rot, np.zeros(8))
    assert np.abs(strain).max() < 1e-15 and np.abs(stress).max() < 1e-13


def test_strain_stress_is_the_elements_virtual_work(fem):
    """The modes recovered by strain_stress are the ones the stiffness and the thermal load condense
    (hex.rs:199-279): on the element's own 2 x 2 x 2 Gauss rule, the virtual work of the stress of
    (u, t) on the strain of du (t = 0) is du . (K u - L t), for random u, t and du. A wrong sign
    or index in the mode recovery breaks it."""
    size, e, nu, alpha = [1.5, 0.8, 1.2], 70.0, 0.33, 2e-3
    k, load = fem.hex_element(size, e, nu, alpha)
    rng = np.random.default_rng(11)
    u, du, t = rng.standard_normal(24), rng.standard_normal(24), rng.standard_normal(8)
    p = 1.0 / np.sqrt(3.0)
    jac = size[0] * size[1] * size[2] / 8.0
    work = 0.0
    for b in VTK_BITS:
        xi = [p if v else -p for v in b]
        _, s = fem.hex_strain_stress(size, e, nu, alpha, xi, u, t)
        ed, _ = fem.hex_strain_stress(size, e, nu, alpha, xi, du, np.zeros(8))
        work += jac * float(s @ ed)
    expect = float(du @ (k @ u - load @ t))
    assert abs(work - expect) <= 1e-11 * (abs(expect) + np.abs(k).max())


def test_isotropic_is_the_orthotropic_special_case(fem):
    """hex.rs:58-64: an isotropic material is the orthotropic one with E, nu and E / (2 (1 + nu)) on
    every axis; the element is the same bit for bit either way."""
    size, e, nu = [1.25, 0.5, 2.0], 210e3, 0.29
    k1 = fem.hex_element(size, e, nu)
    k2 = fem.hex_element(size, [e] * 3, [nu] * 3, g=[e / (2.0 * (1.0 + nu))] * 3)
    assert np.array_equal(k1, k2)


def test_orthotropic_block_held_on_every_face(fem):
    """grid-constrained-block (derivation.md, R. M. Jones 1999 section 2.5): an orthotropic block
    heated by 100 C with every face held normal has no strain, so S sigma = -alpha dT. The stress of
    the element at u = 0 and a uniform dT is -D alpha dT (hex.rs:66-115, 313-340): it equals the
    derivation's values to their 9 decimals, and numpy's solve of the compliance. A swapped Poisson
    ratio or shear modulus fails it."""
    e3, nu3, g3, a3 = [100e3, 50e3, 20e3], [0.3, 0.25, 0.2], [20e3, 10e3, 8e3], [5e-6, 10e-6, 20e-6]
    _, s = fem.hex_strain_stress([2.5, 2.0, 1.5], e3, nu3, a3, [0.1, -0.3, 0.7], np.zeros(24), np.full(8, 100.0), g=g3)
    assert np.allclose(s[:3], [-84.302009777, -72.650733297, -50.027159153], rtol=0, atol=5e-10)
    assert np.all(s[3:] == 0.0)
    ex, ey, ez = e3
    nxy, nxz, nyz = nu3
    comp = np.array([[1 / ex, -nxy / ex, -nxz / ex], [-nxy / ex, 1 / ey, -nyz / ey], [-nxz / ex, -nyz / ey, 1 / ez]])
    assert np.allclose(s[:3], np.linalg.solve(comp, -np.array(a3) * 100.0), rtol=1e-12, atol=0)
    # the shear moduli: a pure engineering shear gamma_xy gives G_xy gamma_xy alone
    x = _corners([2.5, 2.0, 1.5])
    u = np.zeros((8, 3))
    u[:, 0] = 1e-3 * x[:, 1]  # u_x = 1e-3 y
    _, s = fem.hex_strain_stress([2.5, 2.0, 1.5], e3, nu3, a3, [0.0, 0.0, 0.0], u.ravel(), np.zeros(8), g=g3)
    assert np.allclose(s, [0, 0, 0, 0, 0, 20e3 * 1e-3], rtol=1e-12, atol=1e-12)
