"""Assembly on the device (fem_assemble) against the oracle's CPU assembly in the reference's order
(Csr::pattern / add_element, sparse.rs:22-86; submatrix; reduced_rhs, heat.rs:1219-1242): the
reduced matrix and right side equal bit for bit on the box meshes of the three element kinds, on
meshes whose points are moved and renumbered at random, and with a further load per unknown; the
assembled system solves to the oracle's solution."""
import numpy as np
import pytest

from test_assembly import heat_hex8, heat_tet4, mesh_of, tet10

pytestmark = pytest.mark.gpu


def assert_system(a, pr):
    ptr, idx, val, rhs, free = a.arrays()
    rp, ri, rv = pr.reduced().arrays()
    assert np.array_equal(free, pr.free())
    assert np.array_equal(ptr, rp.astype(np.uint64))
    assert np.array_equal(idx, ri)
    assert np.array_equal(val, rv), np.abs(val - rv).max()
    assert np.array_equal(rhs, pr.rhs()), np.abs(rhs - pr.rhs()).max()


@pytest.mark.parametrize("n", [[6, 5, 4], [9, 2, 3]])
def test_heat_hex8_equals_the_reference_assembly_bit_for_bit(fem, oracle, n):
    pr, k = heat_hex8(oracle, n, 1)
    a = fem.Assembly(mesh_of(fem, pr, "heat_hex8", k, source=2.0))
    print(n, a.stats.entries, a.stats.host_seconds, a.stats.device_seconds, flush=True)
    assert_system(a, pr)


@pytest.mark.parametrize("n", [[4, 3, 3], [7, 2, 1]])
def test_heat_tet4_equals_the_reference_assembly_bit_for_bit(fem, oracle, n):
    pr, k = heat_tet4(oracle, n, 2)
    assert_system(fem.Assembly(mesh_of(fem, pr, "heat_tet4", k, source=1.5)), pr)


@pytest.mark.parametrize("n", [[3, 2, 2], [5, 1, 2]])
def test_tet10_equals_the_reference_assembly_bit_for_bit(fem, oracle, n):
    pr, e, nu = tet10(oracle, n, 3)
    assert_system(fem.Assembly(mesh_of(fem, pr, "elastic_tet10", e, nu, body=(0.0, 0.0, -1.0))), pr)


def _moved(pr, corners, rng, amount):
    """The points with the interior corners moved at random (straight tetrahedra stay valid for
    moves below a quarter of a cell), renumbered at random; the cells renumbered and shuffled."""
    xyz = pr.points().copy()
    ijk = pr.ijk()
    top = ijk.max(axis=0)
    inner = np.all((ijk > 0) & (ijk < top), axis=1)
    inner[corners:] = False
    xyz[inner] += rng.uniform(-amount, amount, (int(inner.sum()), 3))
    perm = rng.permutation(xyz.shape[0])
    cells = perm[pr.cells()][rng.permutation(pr.cells().shape[0])]
    return xyz[np.argsort(perm)], cells.astype(np.uint32)


def _reduced_rhs(ptr, idx, val, load, t, fixed):
    """heat.rs:1219-1242 in plain Python floats (IEEE doubles, in the reference's order)."""
    out = []
    for i in np.flatnonzero(~fixed):
        r = float(load[i])
        for e in range(ptr[i], ptr[i + 1]):
            j = idx[e]
            if fixed[j]:
                r -= float(val[e]) * float(t[j])
        out.append(r)
    return np.array(out)


def test_a_moved_renumbered_tet4_mesh_equals_the_reference_assembly(fem, oracle):
    pr, _ = heat_tet4(oracle, [5, 4, 3], 6)
    rng = np.random.default_rng(21)
    h = 0.6 / 3
    xyz, cells = _moved(pr, pr.points().shape[0], rng, 0.2 * h)
    npts = xyz.shape[0]
    k = np.exp(rng.standard_normal(cells.shape[0]))
    fixed = rng.random(npts) < 0.15
    t = np.where(fixed, rng.uniform(-1, 1, npts), 0.0)
    source = 3.0
    ref = oracle.Csr.pattern(npts, cells)
    load = np.zeros(npts)
    for c in range(cells.shape[0]):
        g, v = oracle.tet_gradients(xyz[cells[c]])
        assert v > 0
        ref.add_element(cells[c], oracle.tet_conduction(g, v, [k[c]] * 3))
        for q in range(4):
            load[cells[c][q]] += source * v / 4.0
    fp, fi, fv = ref.arrays()
    rhs = _reduced_rhs(fp, fi, fv, load, t, fixed)
    rp, ri, rv = ref.submatrix(np.flatnonzero(~fixed)).arrays()
    a = fem.Assembly(fem.Mesh("heat_tet4", xyz, cells, k, fixed.astype(np.uint8), t, source=source))
    ptr, idx, val, got, free = a.arrays()
    assert np.array_equal(ptr, rp.astype(np.uint64)) and np.array_equal(idx, ri)
    assert np.array_equal(val, rv)
    assert np.array_equal(got, rhs)


def test_a_moved_tet10_mesh_with_a_further_load_equals_the_reference_assembly(fem, oracle):
    """Straight tet10 (the corners set the element, tet10.rs:71-74); the further load of each
    unknown is added after the body force, as the reference's callers add face loads."""
    pr, e, nu = tet10(oracle, [3, 2, 2], 7)
    rng = np.random.default_rng(22)
    corners = int(pr.sizes.corners)
    xyz, cells = _moved(pr, corners, rng, 0.2)
    npts = xyz.shape[0]
    units = 3 * npts
    fixed = rng.random(units) < 0.1
    t = np.where(fixed, rng.uniform(-0.1, 0.1, units), 0.0)
    extra = rng.standard_normal(units)
    body = (0.5, -0.25, -1.0)
    dofs = (3 * cells[:, :, None] + np.arange(3)[None, None, :]).reshape(cells.shape[0], 30).astype(np.uint32)
    ref = oracle.Csr.pattern(units, dofs)
    load = np.zeros(units)
    for c in range(cells.shape[0]):
        p = xyz[cells[c, :4]]
        normal, shear = oracle.isotropic_stiffness(e[c], nu[c])
        ref.add_element(dofs[c], oracle.tet10_stiffness(p, normal, shear))
        _, v = oracle.tet_gradients(p)
        for i in range(10):
            w = -v / 20.0 if i < 4 else v / 5.0
            for d in range(3):
                load[3 * cells[c, i] + d] += body[d] * w
    load = load + extra
    fp, fi, fv = ref.arrays()
    rhs = _reduced_rhs(fp, fi, fv, load, t, fixed)
    rp, ri, rv = ref.submatrix(np.flatnonzero(~fixed)).arrays()
    a = fem.Assembly(fem.Mesh("elastic_tet10", xyz, cells, e, fixed.astype(np.uint8), t, coef1=nu, body=body,
                              load=extra))
    ptr, idx, val, got, free = a.arrays()
    assert np.array_equal(ptr, rp.astype(np.uint64)) and np.array_equal(idx, ri)
    assert np.array_equal(val, rv), np.abs(val - rv).max()
    assert np.array_equal(got, rhs), np.abs(got - rhs).max()


def test_the_assembled_system_solves_to_the_oracle_solution(fem, oracle):
    """A tet10 box assembled on the device, solved through the assembled boundary (p-multigrid with
    the reference's prolongation), against the oracle's direct solve."""
    pr, e, nu = tet10(oracle, [6, 3, 3], 8)
    a = fem.Assembly(mesh_of(fem, pr, "elastic_tet10", e, nu, body=(0.0, 0.0, -1.0)))
    ptr, idx, val, rhs, free = a.arrays()
    p = pr.prolongation()
    pp, pi, pv = p.arrays()
    s = fem.CsrSolver(ptr, idx, val, "quadratic", p=(pp, pi, pv, p.cols), method="multigrid", tol=1e-12)
    x, st = s.solve(rhs)
    xd, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
