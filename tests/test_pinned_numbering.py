"""Known answers for the integer rules of the boundary, written out by hand from the reference's
rules (no oracle or library code computes the expected values), checked against both the oracle
and the library. CPU only.

- Point numbering of a hex grid: first seen over the cells (x fastest), corners in VTK order
  (crates/fem/src/heat.rs:1165-1208 with grid.rs:169-213).
- Points and tetrahedra of an extruded box: footprint quads row by row, each cut along its
  (0,0)-(1,1) diagonal into (a, b, c) and (a, c, d); per triangle the layers; prism corners first
  seen (prism.rs:244-284); each prism split by split_prism (prism.rs:131-154); a tetrahedron of
  negative volume has its corners 2 and 3 swapped (prism.rs:278-280). The footprint triangulation
  is this library's structured one (the reference meshes footprints with its Delaunay mesher,
  mesh2d.rs, which is not restated); everything after the footprint is the reference's rule.
- Edge nodes of tet10: first seen over the tetrahedra in EDGES order (tet10.rs:38, 244-272),
  numbered after the corners.
- coarsen (multigrid.rs:329-452): kept planes, operator weights lo / (lo + hi), the geometric
  weights next to a fixed plane (where a row has no coupling below), the first-seen coarse
  numbering, rows sorted by coarse index.
"""
import numpy as np
import pytest
import scipy.sparse as sp

# ---------------------------------------------------------------- hex8, 2 x 2 x 1 cells
# Cell 0 numbers its 8 corners 0..7; cell 1 (i=1) meets 4 of them and adds (2,0,0), (2,1,0),
# (2,0,1), (2,1,1); cell 2 (j=1) adds (1,2,0), (0,2,0), (1,2,1), (0,2,1); cell 3 adds (2,2,0), (2,2,1).
HEX_2x2x1_POINTS = [
    (0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1),
    (2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1),
    (1, 2, 0), (0, 2, 0), (1, 2, 1), (0, 2, 1),
    (2, 2, 0), (2, 2, 1),
]
HEX_2x2x1_CELLS = [
    [0, 1, 2, 3, 4, 5, 6, 7],
    [1, 8, 9, 2, 5, 10, 11, 6],
    [3, 2, 12, 13, 7, 6, 14, 15],
    [2, 9, 16, 12, 6, 11, 17, 14],
]

# ---------------------------------------------------------------- tets, 1 x 1 x 1 and 2 x 1 x 1
# 1 x 1 x 1: footprint a=0, b=1, c=3, d=2 (4 footprint points, x fastest). Triangle (0,1,3): global
# prism corners g = [0,1,3,4,5,7], all new: points 0..5. Smallest g at corner 0, min(g1, g5) = 1 <
# min(g2, g4) = 3: tetrahedra [0,1,2,5], [0,1,5,4], [0,4,5,3], all of positive volume. Triangle
# (0,3,2): g = [0,3,2,4,7,6]: new points (0,1,0) -> 6 and (0,1,1) -> 7, nodes [0,2,6,3,5,7];
# min(g1, g5) = 3 > min(g2, g4) = 2: local [0,1,2,4], [0,4,2,5], [0,4,5,3] -> [0,2,6,5], [0,5,6,7],
# [0,5,7,3], all of positive volume.
TET_1x1x1_POINTS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 0), (0, 1, 1)]
TET_1x1x1 = [[0, 1, 2, 5], [0, 1, 5, 4], [0, 4, 5, 3], [0, 2, 6, 5], [0, 5, 6, 7], [0, 5, 7, 3]]
# Kuhn chain of each (axis order of the steps from (0,0,0) to (1,1,1)): xyz, xzy, zxy, yxz, yzx, zyx
TET_1x1x1_TYPES = [0, 1, 4, 2, 3, 5]
# 2 x 1 x 1: the first quad as above; the second quad's triangles (1,2,5) and (1,5,4) bring
# (2,0,0) -> 8, (2,1,0) -> 9, (2,0,1) -> 10, (2,1,1) -> 11; cases A then B as in the first quad.
TET_2x1x1_POINTS = TET_1x1x1_POINTS + [(2, 0, 0), (2, 1, 0), (2, 0, 1), (2, 1, 1)]
TET_2x1x1 = TET_1x1x1 + [[1, 8, 9, 11], [1, 8, 11, 10], [1, 10, 11, 4], [1, 9, 2, 11], [1, 11, 2, 5], [1, 11, 5, 4]]
TET_2x1x1_CELLS = [0] * 6 + [1] * 6
TET_2x1x1_TYPES = TET_1x1x1_TYPES * 2

# ---------------------------------------------------------------- tet10, 1 x 1 x 1
# Edges first seen over TET_1x1x1 in the order 01, 12, 02, 03, 13, 23, numbered from 8.
TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 5), (1, 5), (2, 5), (0, 4), (1, 4), (4, 5), (0, 3), (3, 4), (3, 5),
               (2, 6), (0, 6), (5, 6), (0, 7), (5, 7), (6, 7), (3, 7)]
TET10_1x1x1 = [
    [0, 1, 2, 5, 8, 9, 10, 11, 12, 13],
    [0, 1, 5, 4, 8, 12, 11, 14, 15, 16],
    [0, 4, 5, 3, 14, 16, 11, 17, 18, 19],
    [0, 2, 6, 5, 10, 20, 21, 11, 13, 22],
    [0, 5, 6, 7, 11, 22, 21, 23, 24, 25],
    [0, 5, 7, 3, 11, 24, 23, 17, 19, 26],
]


def test_hex_grid_points_are_first_seen_in_vtk_order(fem, oracle):
    n = [2, 2, 1]
    pr = oracle.Problem.heat(n, [1.0, 1.0, 0.5], np.ones(4), fixed={"x0": 0.0})
    assert pr.ijk().tolist() == [list(p) for p in HEX_2x2x1_POINTS]
    assert pr.cells().tolist() == HEX_2x2x1_CELLS
    m = fem.Model(fem.Box("heat_hex8", n, [1.0, 1.0, 0.5], np.ones(4), fixed={"x0": 0.0}))
    assert m.point_grid().tolist() == [list(p) for p in HEX_2x2x1_POINTS]
    # free list: the points off x = 0, in increasing order of number (heat.rs:503)
    assert m.free_list().tolist() == [k for k, p in enumerate(HEX_2x2x1_POINTS) if p[0] != 0]


@pytest.mark.parametrize("n,points,tets,cells,types", [
    ([1, 1, 1], TET_1x1x1_POINTS, TET_1x1x1, [0] * 6, TET_1x1x1_TYPES),
    ([2, 1, 1], TET_2x1x1_POINTS, TET_2x1x1, TET_2x1x1_CELLS, TET_2x1x1_TYPES),
])
def test_extruded_box_points_and_tetrahedra(fem, oracle, n, points, tets, cells, types):
    nt = 6 * int(np.prod(n))
    pr = oracle.Problem.heat(n, [float(n[0]), 1.0, 1.0], np.ones(nt), fixed={"x0": 0.0}, tet4=True)
    assert pr.ijk().tolist() == [list(p) for p in points]
    assert pr.cells().tolist() == tets
    m = fem.Model(fem.Box("heat_tet4", n, [float(n[0]), 1.0, 1.0], np.ones(nt), fixed={"x0": 0.0}))
    assert m.point_grid().tolist() == [list(p) for p in points]
    cell, kind = m.tet_cells()
    assert cell.tolist() == cells
    assert kind.tolist() == types


def test_tet10_edge_nodes_are_first_seen_in_edges_order(fem, oracle):
    pr = oracle.Problem.elastic(10, [1, 1, 1], [1.0, 1.0, 1.0], np.ones(6), np.full(6, 0.3), clamped=["x0"])
    assert pr.cells().tolist() == TET10_1x1x1
    assert pr.edges().tolist() == [list(e) for e in TET10_EDGES]
    # points on the grid refined twice: corners at twice their cell coordinates, edge nodes at the
    # sum of their corners' coordinates
    want = [[2 * c for c in p] for p in TET_1x1x1_POINTS]
    want += [[TET_1x1x1_POINTS[a][d] + TET_1x1x1_POINTS[b][d] for d in range(3)] for a, b in TET10_EDGES]
    assert pr.ijk().tolist() == want
    m = fem.Model(fem.Box("elastic_tet10", [1, 1, 1], [1.0, 1.0, 1.0], 1.0, 0.3, fixed={"x0": True}))
    assert m.point_grid().tolist() == want


# ---------------------------------------------------------------- coarsen
# Heat on 4 x 2 x 2 unit hex cells (5 x 3 x 3 points), conductivity 1, 1, 3, 1 by x layer, x = 0
# fixed; the 36 free points in the order x fastest (the test permutes the reduced matrix to it).
# Kept planes: x 0, 2, 4 (coarse 0, 1, 2), y and z 0, 2 (coarse 0, 1).
# Weights per axis of a fine point (coarse plane, weight):
#   x = 1: no coupling to the fixed plane below, so the geometric (0, 1/2), (1, 1/2)
#   x = 3: lo from the k = 3 layer, hi from the k = 1 layer: (1, 3/4), (2, 1/4)
#   x = 2, 4: (1, 1), (2, 1)
#   y = 1 (and z = 1): equal couplings below and above: (0, 1/2), (1, 1/2)
# First-seen coarse numbering over the fine rows, x outermost in each row's tensor loop:
COARSE_POINTS = [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 0), (2, 1, 0),
                 (0, 0, 1), (1, 0, 1), (2, 0, 1), (0, 1, 1), (1, 1, 1), (2, 1, 1)]
X_WEIGHTS = {1: [(0, 0.5), (1, 0.5)], 2: [(1, 1.0)], 3: [(1, 0.75), (2, 0.25)], 4: [(2, 1.0)]}
YZ_WEIGHTS = {0: [(0, 1.0)], 1: [(0, 0.5), (1, 0.5)], 2: [(1, 1.0)]}
# Four rows written out in full: fine point -> [(coarse index, weight)] sorted by index.
ROWS = {
    (1, 0, 0): [(0, 0.5), (1, 0.5)],
    (3, 1, 1): [(1, 3 / 16), (2, 1 / 16), (4, 3 / 16), (5, 1 / 16), (7, 3 / 16), (8, 1 / 16), (10, 3 / 16), (11, 1 / 16)],
    (1, 1, 2): [(6, 0.25), (7, 0.25), (9, 0.25), (10, 0.25)],
    (4, 2, 2): [(11, 1.0)],
}


def _coarsen_input(oracle):
    n = [4, 2, 2]
    kx = np.array([1.0, 1.0, 3.0, 1.0])
    k = np.array([kx[c % 4] for c in range(16)])
    pr = oracle.Problem.heat(n, [4.0, 2.0, 2.0], k, fixed={"x0": 0.0})
    a = pr.reduced().to_scipy().tocsr()
    pts = pr.ijk()[pr.free()]
    order = np.argsort(pts[:, 0] + 5 * (pts[:, 1] + 3 * pts[:, 2]), kind="stable")
    a = sp.csr_matrix(a[order][:, order])
    a.sort_indices()
    planes = [np.arange(5, dtype=np.float64), np.arange(3, dtype=np.float64), np.arange(3, dtype=np.float64)]
    return a, pts[order].astype(np.uint32), planes


def _check_prolongation(ptr, idx, val, fine_points):
    ptr = np.asarray(ptr, np.int64)
    for r, (i, j, k) in enumerate(fine_points.tolist()):
        cols, vals = idx[ptr[r]:ptr[r + 1]].tolist(), val[ptr[r]:ptr[r + 1]]
        assert cols == sorted(cols)
        # every entry is the product of the hand-written axis weights of its coarse point
        want = {}
        for cx, wx in X_WEIGHTS[i]:
            for cy, wy in YZ_WEIGHTS[j]:
                for cz, wz in YZ_WEIGHTS[k]:
                    want[COARSE_POINTS.index((cx, cy, cz))] = wx * wy * wz
        assert cols == sorted(want)
        assert np.allclose(vals, [want[c] for c in cols], rtol=0, atol=1e-15)
        if (i, j, k) in ROWS:
            assert cols == [c for c, _ in ROWS[(i, j, k)]]
            assert np.allclose(vals, [w for _, w in ROWS[(i, j, k)]], rtol=0, atol=1e-15)


def test_coarsen_known_answer_oracle(oracle):
    a, pts, planes = _coarsen_input(oracle)
    m = oracle.Csr.from_arrays(a.shape[0], a.indptr, a.indices, a.data)
    p, cpts, ccomp = oracle.coarsen(m, planes, pts, np.zeros(len(pts), np.uint8))
    assert p.rows == 36 and p.cols == 12
    assert cpts.tolist() == [list(q) for q in COARSE_POINTS]
    ptr, idx, val = p.arrays()
    _check_prolongation(ptr, idx, val, pts)


def test_coarsen_known_answer_library(fem, oracle):
    a, pts, planes = _coarsen_input(oracle)
    levels = fem.csr_levels(a.indptr.astype(np.uint64), a.indices.astype(np.uint32), a.data,
                            fem.Grid(planes, pts, np.zeros(len(pts), np.uint8)), coarsest=1, max_levels=2)
    assert len(levels) == 2
    assert levels[1]["points"].tolist() == [list(q) for q in COARSE_POINTS]
    pp, pi, pv, pc = levels[0]["p"]
    assert pc == 12
    _check_prolongation(pp, pi, pv, pts)
