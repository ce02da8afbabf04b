"""The library's host-side model (numbering, free unknowns, loads) against the oracle, with no
device: integer structures bit-exact, loads to rounding. CPU only."""
import numpy as np
import pytest

CASES = [
    ("heat_hex8", [1, 1, 1]), ("heat_hex8", [5, 1, 3]), ("heat_hex8", [4, 3, 2]),
    ("heat_tet4", [1, 1, 1]), ("heat_tet4", [3, 4, 2]), ("heat_tet4", [2, 1, 5]),
    ("elastic_tet4", [1, 1, 1]), ("elastic_tet4", [4, 2, 3]),
    ("elastic_tet10", [1, 1, 1]), ("elastic_tet10", [3, 2, 2]), ("elastic_tet10", [2, 5, 1]),
]
FIX_SETS = [{"x0": 0.0}, {"x0": 1.5, "x1": -0.5}, {"y0": 2.0, "z1": 1.0, "x1": 0.0}]


def build_pair(fem, oracle, kind, n, variant):
    rng = np.random.default_rng(100 + variant)
    length = [1.0 + 0.3 * variant, 0.7, 1.3]
    cells = int(np.prod(n))
    if kind.startswith("heat"):
        tet4 = kind == "heat_tet4"
        k = rng.uniform(0.2, 3.0, cells * (6 if tet4 else 1))
        fixed = FIX_SETS[variant % len(FIX_SETS)]
        src = 0.25 * variant
        pr = oracle.Problem.heat(n, length, k, fixed=fixed, source=src, tet4=tet4)
        box = fem.Box(kind, n, length, k, fixed=fixed, source=src)
    else:
        order = 4 if kind == "elastic_tet4" else 10
        e = rng.uniform(1.0, 3.0, 6 * cells)
        nu = rng.uniform(0.1, 0.45, 6 * cells)
        clamped = [["x0"], ["x0", "z1"], ["y1"]][variant % 3]
        body = [(0, 0, -1), (0.3, -0.2, 0.1), (0, 0, 0)][variant % 3]
        traction = [{"x1": (0, 0, -0.5)}, {"z0": (1, 2, 3), "y1": (0, -1, 0)}, {}][variant % 3]
        pr = oracle.Problem.elastic(order, n, length, e, nu, clamped=clamped, body=body, traction=traction)
        box = fem.Box(kind, n, length, e, nu, fixed={f: True for f in clamped}, body=body, traction=traction)
    return pr, fem.Model(box)


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("kind,n", CASES)
def test_numbering_free_list_and_loads_match_the_oracle(fem, oracle, kind, n, variant):
    pr, m = build_pair(fem, oracle, kind, n, variant)
    assert m.points() == pr.sizes.n_points
    assert np.array_equal(m.point_grid(), pr.ijk())
    assert np.array_equal(m.free_list(), pr.free())
    want = pr.load()
    got = m.load()
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max())


def test_each_tet10_point_sits_on_the_refined_grid(fem):
    """Corners at even refined coordinates, edge nodes at the midpoint of their edge."""
    n = [3, 2, 2]
    m = fem.Model(fem.Box("elastic_tet10", n, [1, 1, 1], 1.0, 0.3, fixed={"x0": True}))
    g = m.point_grid()
    corners = (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
    assert m.points() == (2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1)
    assert (g[:corners] % 2 == 0).all()
    assert (g[corners:] % 2 == 1).any(axis=1).all()
    assert len({tuple(p) for p in g.tolist()}) == m.points()


def test_material_errors_use_the_reference_texts(fem):
    with pytest.raises(fem.FemError) as e:
        fem.Model(fem.Box("elastic_tet4", [1, 1, 1], [1, 1, 1], 1.0, 0.5, fixed={"x0": True}))
    assert "the Poisson ratios give no stable material" in str(e.value)  # hex.rs:93
    with pytest.raises(fem.FemError) as e:
        fem.Model(fem.Box("elastic_tet4", [1, 1, 1], [1, 1, 1], -1.0, 0.3, fixed={"x0": True}))
    assert "the moduli must be above 0 and the Poisson ratios finite" in str(e.value)  # hex.rs:79


@pytest.mark.parametrize("kind,n", [("heat_tet4", [3, 4, 2]), ("elastic_tet10", [2, 5, 1]), ("elastic_tet4", [4, 2, 3])])
def test_tet_centroids_follow_the_reference_numbering(fem, oracle, kind, n):
    """Model.tet_centroids (from fem_model_tet_cells) against the centroids of the oracle's own
    tetrahedra, which it numbers by the reference's extrusion and prism split."""
    length = [1.3, 0.7, 2.1]
    nt = 6 * int(np.prod(n))
    if kind == "heat_tet4":
        pr = oracle.Problem.heat(n, length, np.ones(nt), fixed={"x0": 0.0}, tet4=True)
        m = fem.Model(fem.Box(kind, n, length, np.ones(nt), fixed={"x0": 0.0}))
    else:
        order = 4 if kind == "elastic_tet4" else 10
        pr = oracle.Problem.elastic(order, n, length, np.ones(nt), np.full(nt, 0.3), clamped=["x0"])
        m = fem.Model(fem.Box(kind, n, length, 1.0, 0.3, fixed={"x0": True}))
    want = pr.points()[pr.cells()[:, :4]].mean(axis=1)
    assert np.abs(m.tet_centroids() - want).max() <= 1e-14
    cell, kind_ = m.tet_cells()
    assert np.array_equal(np.bincount(cell, minlength=int(np.prod(n))), np.full(int(np.prod(n)), 6))
    assert set(np.unique(kind_).tolist()) == set(range(6))
