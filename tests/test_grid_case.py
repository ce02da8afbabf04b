"""The grid.toml front end on the host (no device): reading the reference's case files, the graded
planes (crates/fem/src/grid.rs:411-521), part assignment, symmetry crop (heat.rs:1483-1573),
numbering (heat.rs:1165-1208) and probe weights.

The graded planes are checked against the rules they must obey (every part face on a plane, no
cell longer than max_cell, at least min_cells through each part, neighbours within max_growth),
and the numbering against the library's own box model (model.cpp) on a one-part box: two
restatements of heat.rs:1165-1208 that must agree.
"""
import os

import numpy as np
import pytest

from femahoy_amd import grid_case as gc

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")


def case(path):
    return gc.load_case(os.path.join(CASES, path, "grid.toml"))


def built(c):
    used = [any(l.get("faces") == f for l in c.loads) for f in range(len(c.faces))]
    return gc.build_grid(c, used, [(0.0,) for _ in c.parts])


def test_the_four_cases_read():
    bar, beam, wall, block = (case(p) for p in ("grid-heated-bar/bar", "grid-cantilever/beam", "grid-layered-wall/wall",
                                                "grid-heated-block-ccx/block"))
    assert (bar.step, beam.step, wall.step, block.step) == ("steady", "static", "steady", "steady")
    assert bar.symmetry == [False, True, True] and bar.max_cell == 0.000625 and bar.min_cells == 2
    assert [l["kind"] for l in beam.loads] == ["displacement", "traction"]
    assert beam.loads[0]["value"] == [0.0, 0.0, 0.0] and beam.loads[1]["value"] == [0.0, 0.0, -1.0]
    assert [p.name for p in wall.parts] == ["plaster", "insulation", "brick"] and wall.min_cells == 2
    assert [l["kind"] for l in wall.loads] == ["convection", "convection"]
    assert [n for n, _ in block.probes] == ["block_top", "plate_corner", "plate_edge_top"]


@pytest.mark.parametrize("text,msg", [
    ('units = "SI-m"\n[grid]\nmax_cell = 1.0\n[material.a]\nconductivity = 1.0\nexpansion = 1e-6\n'
     '[part.p]\nmaterial = "a"\nx = [0.0, 1.0]\ny = [0.0, 1.0]\nz = [0.0, 1.0]\n[step]\ntype = "steady"\n',
     "`expansion` is not read"),
    ('units = "SI-m"\n[grid]\nmax_cell = 1.0\n[material.a]\nyoungs_modulus = 1.0\npoisson_ratio = 0.3\n'
     '[part.p]\nmaterial = "a"\nx = [0.0, 1.0]\ny = [0.0, 1.0]\nz = [0.0, 1.0]\n[step]\ntype = "modes"\n',
     "this front end solves steady and transient heat and static stress"),
    ('units = "SI-cm"\n[grid]\nmax_cell = 1.0\n[step]\ntype = "steady"\n', "`units` must be SI-m or SI-mm"),
    ('units = "SI-m"\n[grid]\nmax_cell = 1.0\n[material.a]\nconductivity = 1.0\n'
     '[part.p]\nmaterial = "b"\nx = [0.0, 1.0]\ny = [0.0, 1.0]\nz = [0.0, 1.0]\n[step]\ntype = "steady"\n',
     "no material b"),
])
def test_keys_outside_the_front_end_are_refused(text, msg):
    with pytest.raises(gc.CaseError, match=msg):
        gc.read(text)


def test_graded_planes_obey_the_grid_rules():
    c = case("grid-layered-wall/wall")
    g, _, _ = built(c)
    x = np.array(g.planes[0])
    for p in c.parts:
        for v in (p.lo[0], p.hi[0]):
            assert np.min(np.abs(x - v)) <= 1e-12
        inside = np.sum((x[:-1] >= p.lo[0] - 1e-12) & (x[1:] <= p.hi[0] + 1e-12))
        assert inside >= c.min_cells
    h = np.diff(x)
    assert np.all(h <= c.max_cell * (1 + 1e-12))
    ratio = np.maximum(h[1:] / h[:-1], h[:-1] / h[1:])
    assert np.all(ratio <= c.max_growth * (1 + 1e-9))
    assert not np.allclose(h, h[0])  # graded: the plaster's cells are smaller than the brick's
    # the 3 materials in x order
    mats = [c.parts[int(g.part[g.cell_index([i, 0, 0])])].name for i in range(len(h))]
    assert mats[0] == "plaster" and mats[-1] == "brick" and "insulation" in mats


def test_uniform_axes_are_equal_steps():
    g, _, _ = built(case("grid-cantilever/beam"))
    assert [len(p) - 1 for p in g.planes] == [40, 4, 4]
    assert np.allclose(np.diff(g.planes[0]), 2.5, rtol=0, atol=1e-12)


def test_symmetry_keeps_the_part_above_the_mid_planes():
    g, faces, mids = built(case("grid-heated-bar/bar"))
    assert mids == [None, 0.01, 0.01]
    assert [len(p) - 1 for p in g.planes] == [4, 16, 16]
    assert g.planes[1][0] == pytest.approx(0.01) and g.planes[2][-1] == pytest.approx(0.02)
    # the fixed sides kept are y+ and z+ only: 4 x 16 cell faces each
    assert sorted({s for _, s in faces[0]}) == [3, 5] and len(faces[0]) == 128


def test_void_cells_and_their_faces():
    g, faces, mids = built(case("grid-heated-block-ccx/block"))
    assert mids[0] == 0.05 and mids[1] == 0.05
    void = int(np.sum(g.part == gc.VOID))
    assert void > 0 and void < len(g.part)
    # the convection face set (z- of the plate) covers the plate's whole bottom after the crop
    nx, ny, _ = g.dims()
    assert len(faces[0]) == nx * ny and {s for _, s in faces[0]} == {4}


def test_numbering_equals_the_library_box_model(fem):
    text = ('units = "SI-m"\n[grid]\nmax_cell = 0.25\n[material.a]\nconductivity = 1.0\n'
            '[part.p]\nmaterial = "a"\nx = [0.0, 1.5]\ny = [0.0, 0.75]\nz = [0.0, 1.0]\n'
            '[face.f]\nside = "x-"\n[[load]]\ntemperature = 0.0\nface = "f"\n[step]\ntype = "steady"\n')
    c = gc.read(text)
    g, _, _ = built(c)
    ijk, cells, cell_of, _ = gc.number(g)
    n = [len(p) - 1 for p in g.planes]
    assert n == [6, 3, 4]
    model = fem.Model(fem.Box("heat_hex8", n, [1.5, 0.75, 1.0], 1.0, fixed={"x0": 0.0}))
    assert np.array_equal(ijk, model.point_grid())
    assert cell_of == list(range(int(np.prod(n))))


def test_probe_at_a_grid_point_reads_that_point():
    c = case("grid-cantilever/beam")
    g, _, mids = built(c)
    _, _, _, number_of = gc.number(g)
    w = {name: pairs for name, pairs, *_ in gc.probe_weights(c, g, mids, number_of)}
    # the graded planes are placed by the size function's integral, so a plane is at its
    # coordinate to rounding: the weights are 1 and 0 to 1e-12
    tip = {v: wt for v, wt in w["tip"] if abs(wt) > 1e-12}
    assert list(tip) == [number_of[(40, 2, 2)]] and abs(tip[number_of[(40, 2, 2)]] - 1.0) <= 1e-12
    # middle_top at x = 51.25 is halfway between planes 20 and 21: 1/2 on each corner of the top edge
    mid = sorted(wt for _, wt in w["middle_top"] if abs(wt) > 1e-12)
    assert np.allclose(mid, [0.5, 0.5], rtol=0, atol=1e-12)


def test_static_thermal_case_reads():
    """grid-heated-block-stress-ccx: heat loads in a static step are the temperature of its thermal
    strain (core grid.rs:814-850), a point support holds uz at the centre of the bottom, and the
    bottom face set is a surface whose shape the result gives."""
    c = case("grid-heated-block-stress-ccx/block")
    assert c.step == "static" and c.thermal == {"reference": 20.0, "heat": True}
    assert [(l["kind"], bool(l.get("heat"))) for l in c.loads] == [("convection", True), ("source", True),
                                                                   ("point", False)]
    assert c.loads[2]["point"] == [0.05, 0.05, 0.0] and c.loads[2]["value"] == [None, None, 0.0]
    assert [m["expansion"] for m in c.materials] == [[12e-6] * 3, [17e-6] * 3]
    assert c.surfaces == [("bottom", 0)]
    g, _, mids = built(c)
    ijk, _, _, number_of = gc.number(g)
    node = gc.grid_point(g, number_of, mids, [0.05, 0.05, 0.0])
    assert tuple(ijk[node]) == (0, 0, 0)
    # a point below a plane of symmetry reads its mirror point
    assert gc.grid_point(g, number_of, mids, [0.0, 0.0, 0.0]) == number_of[(len(g.planes[0]) - 1,
                                                                             len(g.planes[1]) - 1, 0)]
    with pytest.raises(gc.CaseError, match="is not a point of the grid of the parts"):
        gc.grid_point(g, number_of, mids, [0.0501, 0.05, 0.0])


BLOCK = os.path.join(CASES, "grid-heated-block-stress-ccx", "block", "grid.toml")


@pytest.mark.parametrize("edit,msg", [
    (("reference_temperature = 20.0\n", ""), "missing key `reference_temperature`"),
    (('type = "static"\n', 'type = "static"\ntemperature = 80.0\n'), "give `temperature` or heat loads"),
    (("expansion = 17e-6\n", "expansion = 17e-6\n"), None),
    (('[surface.bottom]\nface = "bottom"\n', '[surface.bottom]\nface = "top"\n'), "no face set top"),
    (("conductivity = 400.0\n", ""), "need the conductivity of every material"),
])
def test_static_thermal_errors(edit, msg):
    with open(BLOCK) as f:
        text = f.read()
    assert edit[0] in text
    text = text.replace(edit[0], edit[1])
    if msg is None:
        gc.read(text)
    else:
        with pytest.raises(gc.CaseError, match=msg):
            gc.read(text)


def test_reference_temperature_needs_a_thermal_load():
    text = open(os.path.join(CASES, "grid-cantilever", "beam", "grid.toml")).read()
    text = text.replace('type = "static"', 'type = "static"\nreference_temperature = 20.0')
    with pytest.raises(gc.CaseError, match="`reference_temperature` needs `temperature` or heat loads"):
        gc.read(text)
    uniform = gc.read(text.replace("reference_temperature = 20.0", "reference_temperature = 20.0\ntemperature = 70.0"))
    assert uniform.thermal == {"reference": 20.0, "uniform": 70.0}


def test_solve_small_is_gauss_elimination():
    """elastic.rs:1252-1286 restated: a regular system equals numpy's solve; a singular one gives
    0 for the unknowns with no pivot and solves the rest."""
    rng = np.random.default_rng(3)
    b = rng.standard_normal((6, 6))
    m = b @ b.T + 6 * np.eye(6)
    r = rng.standard_normal(6)
    for k in (3, 6):
        x = gc._solve_small(m.tolist(), r.tolist(), k)
        assert np.allclose(x[:k], np.linalg.solve(m[:k, :k], r[:k]), rtol=1e-12, atol=0)
        assert x[k:] == [0.0] * (6 - k)
    # samples on one line a = const: the column of s is zero, its unknown stays 0
    m2 = np.diag([2.0, 0.0, 3.0, 1.0, 1.0, 1.0])
    x = gc._solve_small(m2.tolist(), [4.0, 5.0, 6.0, 0, 0, 0], 3)
    assert x[:3] == [2.0, 0.0, 2.0]


def _shape_of(w_of_xy):
    c = case("grid-heated-block-stress-ccx/block")
    g, faces, mids = gc.build_grid(c, [True], [(0.0, 0.0, 0.0)] * len(c.parts))
    ijk, _, _, number_of = gc.number(g)
    u = np.zeros(3 * len(ijk))
    for (i, j, k), v in number_of.items():
        # the bottom face is z-: w = -uz
        u[3 * v + 2] = -w_of_xy(g.planes[0][i], g.planes[1][j])
    return gc.surface_shape(g, faces[0], mids, number_of, u, "bottom")


def test_surface_shape_of_a_plane_and_a_bowl():
    """elastic.rs:1099-1250 on the stress block's bottom face (a quarter, mirrored across x and y):
    w mirrors with the faces, so a field of the quarter stands for an even field of the whole face.
    A uniform w has no range, flatness or curvature; w = k/2 (x - 0.05)^2 + k/4 (y - 0.05)^2 has the
    curvatures -k and -k/2 and no twist (to the error of the bilinear faces, h^2 / L^2), and its
    range is the largest minus the smallest corner value. A quarter's tilted plane w = c (x - 0.05)
    is the whole face's V, w = c |x - 0.05|: its least squares plane is flat (its mean), so its
    flatness equals its range."""
    s = _shape_of(lambda x, y: 1e-6)
    assert s["bottom.range"] == 0.0 and abs(s["bottom.flatness"]) < 1e-18
    assert max(abs(s[k]) for k in ("bottom.kxx", "bottom.kyy", "bottom.kxy")) < 1e-12
    s = _shape_of(lambda x, y: 2e-5 * (x - 0.05))
    assert abs(s["bottom.range"] - 1e-6) < 1e-18 and abs(s["bottom.flatness"] - 1e-6) < 1e-15
    k = 3e-3
    s = _shape_of(lambda x, y: 0.5 * k * (x - 0.05) ** 2 + 0.25 * k * (y - 0.05) ** 2)
    assert abs(s["bottom.kxx"] + k) < 2e-3 * k and abs(s["bottom.kyy"] + 0.5 * k) < 2e-3 * k
    assert abs(s["bottom.kxy"]) < 1e-9 * k
    assert s["bottom.k1"] >= s["bottom.k2"] and s["bottom.shape"] == "concave"
    assert abs(s["bottom.range"] - (0.5 * k + 0.25 * k) * 0.05 ** 2) < 1e-15
