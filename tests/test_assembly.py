"""Assembly on the device, host part (fem_assemble's pattern, free list and reduced pattern: the
reference's Csr::pattern, sparse.rs:22-62, and submatrix, sparse.rs:108-134) against the oracle's
bit for bit, on the box meshes of the three element kinds and on a mesh whose points and cells are
renumbered at random; and the error texts. CPU only (no device call)."""
import numpy as np
import pytest


def mesh_of(fem, pr, kind, coef0, coef1=None, source=0.0, body=(0.0, 0.0, 0.0)):
    fixed, value = pr.fixed()
    return fem.Mesh(kind, pr.points(), pr.cells(), coef0, fixed, value, coef1=coef1, source=source, body=body)


def heat_hex8(oracle, n, seed):
    k = np.exp(np.random.default_rng(seed).standard_normal(int(np.prod(n))))
    return oracle.Problem.heat(n, [1.0, 0.8, 0.6], k, fixed={"x0": 0.0, "x1": 1.0}, source=2.0), k


def heat_tet4(oracle, n, seed):
    k = np.exp(np.random.default_rng(seed).standard_normal(6 * int(np.prod(n))))
    return oracle.Problem.heat(n, [1.0, 0.8, 0.6], k, fixed={"x0": 0.0, "y1": 1.0}, source=1.5, tet4=True), k


def tet10(oracle, n, seed):
    rng = np.random.default_rng(seed)
    nt = 6 * int(np.prod(n))
    e, nu = rng.uniform(1.0, 3.0, nt), rng.uniform(0.2, 0.4, nt)
    return oracle.Problem.elastic(10, n, [float(v) for v in n], e, nu, clamped=["x0"], body=(0.0, 0.0, -1.0)), e, nu


def assert_pattern(fem, mesh, ref):
    ptr, idx, free = fem.assembly_pattern(mesh)
    rp, ri, _ = ref.reduced().arrays()
    assert np.array_equal(free, ref.free())
    assert np.array_equal(ptr, rp.astype(np.uint64))
    assert np.array_equal(idx, ri)


@pytest.mark.parametrize("n", [[6, 5, 4], [3, 1, 7]])
def test_the_heat_hex8_pattern_is_the_reference_pattern(fem, oracle, n):
    pr, k = heat_hex8(oracle, n, 1)
    assert_pattern(fem, mesh_of(fem, pr, "heat_hex8", k, source=2.0), pr)


@pytest.mark.parametrize("n", [[4, 3, 3], [5, 2, 1]])
def test_the_heat_tet4_pattern_is_the_reference_pattern(fem, oracle, n):
    pr, k = heat_tet4(oracle, n, 2)
    assert_pattern(fem, mesh_of(fem, pr, "heat_tet4", k, source=1.5), pr)


@pytest.mark.parametrize("n", [[3, 2, 2], [4, 1, 1]])
def test_the_tet10_pattern_is_the_reference_pattern(fem, oracle, n):
    pr, e, nu = tet10(oracle, n, 3)
    assert_pattern(fem, mesh_of(fem, pr, "elastic_tet10", e, nu, body=(0.0, 0.0, -1.0)), pr)


def test_a_renumbered_mesh_has_the_pattern_of_csr_pattern(fem, oracle):
    """Points and cells in a random order (an unstructured numbering): the reduced pattern equals
    the oracle's Csr::pattern of the same cells, restricted to the free unknowns."""
    pr, k = heat_tet4(oracle, [4, 3, 2], 5)
    rng = np.random.default_rng(11)
    npts = pr.points().shape[0]
    perm = rng.permutation(npts)          # new number of old point p
    inv = np.argsort(perm)
    cells = perm[pr.cells()][rng.permutation(pr.cells().shape[0])].astype(np.uint32)
    xyz = pr.points()[inv]
    fixed = rng.random(npts) < 0.2
    mesh = fem.Mesh("heat_tet4", xyz, cells, np.ones(cells.shape[0]), fixed.astype(np.uint8), np.zeros(npts))
    ptr, idx, free = fem.assembly_pattern(mesh)
    ref = oracle.Csr.pattern(npts, cells).submatrix(np.flatnonzero(~fixed))
    rp, ri, _ = ref.arrays()
    assert np.array_equal(free, np.flatnonzero(~fixed))
    assert np.array_equal(ptr, rp.astype(np.uint64))
    assert np.array_equal(idx, ri)


def test_bad_input_gives_a_message(fem, oracle):
    pr, e, nu = tet10(oracle, [2, 1, 1], 4)
    bad = mesh_of(fem, pr, "elastic_tet10", e, np.full_like(nu, 0.5))
    with pytest.raises(fem.FemError, match="^the Poisson ratios give no stable material"):
        fem.Assembly(bad)
    bad = mesh_of(fem, pr, "elastic_tet10", -e, nu)
    with pytest.raises(fem.FemError, match="^the moduli must be above 0"):
        fem.Assembly(bad)
    cells = pr.cells().copy()
    cells[0, 0] = pr.points().shape[0]
    fixed, value = pr.fixed()
    with pytest.raises(fem.FemError, match="names a point that the mesh does not have"):
        fem.assembly_pattern(fem.Mesh("elastic_tet10", pr.points(), cells, e, fixed, value, coef1=nu))
