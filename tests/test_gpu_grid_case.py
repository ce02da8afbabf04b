"""The reference's own validation cases, read from their grid.toml files and solved on the device
through the grid.toml front end (femahoy_amd.grid_case): each probe passes the case's check.toml
with the reference's pass rule (|computed - reference| <= rel_tol |reference| + abs_tol,
crates/fh/src/validation/run.rs:38), and where a case's derivation notes print what the
reference's own grid solver gave on this mesh, the probe matches it to half a unit of its last
printed digit.

The case files are fixtures copied from the reference's validation/cases (grid.toml and
check.toml as they are): a heated square bar with quarter symmetry (grid-heated-bar), a hex
cantilever with incompatible modes (grid-cantilever), a wall of 3 layers with a convection on
each side on a graded grid (grid-layered-wall), and a steel plate with a heated copper block on
it, void cells beside the block and half symmetry in x and y (grid-heated-block-ccx), and the
same plate and block with the thermal stress of its steady temperature, a point support and the
shape of the bottom surface (grid-heated-block-stress-ccx). The transient cooling slab
(grid-cooling-slab) and the uniformly heated bimetal strip (grid-bimetal-strip) are checked against
their derivations' whole tables of grids.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")

# derivation.md of each case: the reference grid solver's value on the case's own mesh
PRINTED = {
    ("grid-heated-bar", "centre"): (14.745623, 0.5e-6),      # max_cell 0.625 mm
    ("grid-cantilever", "tip.uz"): (-0.190113, 0.5e-6),      # max_cell 2.5 mm
    ("grid-cantilever", "middle_top.sxx"): (29.25, 0.5e-7),   # "the grid gives 29.2500000 MPa"
    # max_cell 1.25 mm, 52,344 unknowns, values in 1e-6 m to 4 decimals
    ("grid-heated-block-stress-ccx", "block_top.uz"): (12.7391e-6, 0.5e-10),
    ("grid-heated-block-stress-ccx", "plate_corner.ux"): (10.2845e-6, 0.5e-10),
    ("grid-heated-block-stress-ccx", "plate_corner.uz"): (-7.2154e-6, 0.5e-10),
    ("grid-heated-block-stress-ccx", "plate_edge.uz"): (-4.8369e-6, 0.5e-10),
}


@pytest.mark.parametrize("case_dir", ["grid-heated-bar/bar", "grid-cantilever/beam", "grid-layered-wall/wall",
                                      "grid-heated-block-ccx/block", "grid-heated-block-stress-ccx/block",
                                      "grid-constrained-block/block"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_reference_case_passes_its_check(fem, case_dir, precision):
    from femahoy_amd import grid_case
    folder = case_dir.split("/")[0]
    case = grid_case.load_case(os.path.join(CASES, case_dir, "grid.toml"))
    res = grid_case.solve(case, precision=precision, tol=1e-12)
    print(case_dir, precision, res.method, res.unknowns, res.stats)
    with open(os.path.join(CASES, folder, "check.toml")) as f:
        rows = grid_case.check(res, f.read(), case.name)
    assert rows, "no check of this case reads a probe"
    for key, column, got, ref, limit, ok in rows:
        print(f"  {key}: {column} = {got!r} (reference {ref}, limit {limit:.3g})")
        assert ok, (key, got, ref, limit)
    # every entry of the check that names this case is computed (probes, stresses, surfaces)
    doc = grid_case._toml.loads(open(os.path.join(CASES, folder, "check.toml")).read())
    assert len(rows) == sum(1 for v in doc.values() if isinstance(v, dict) and v.get("case") == case.name)
    for (c, column), (value, half) in PRINTED.items():
        if c == folder:
            assert abs(res.probes[column] - value) <= half, (column, res.probes[column], value)
    if folder == "grid-heated-block-stress-ccx":
        # the centre of the bottom holds uz = 0 and the corner is its lowest point: the range of
        # the bottom surface is -plate_corner.uz, and the plate bends down at its edges (concave)
        assert res.unknowns == 52344
        assert abs(res.probes["bottom.range"] + res.probes["plate_corner.uz"]) <= 1e-12 * res.probes["bottom.range"]
        assert res.probes["bottom.shape"] == "concave"


# grid-cooling-slab (derivation.md, "Discretization error and tolerance"): the reference's grid solver's
# centre temperature at t = 20 s for 3 grids and 4 time steps (backward Euler), printed to 5 decimals.
COOLING = {0.001: (77.16018, 77.13918, 77.12847, 77.12306),
           0.0005: (77.24674, 77.22512, 77.21410, 77.20853),
           0.00025: (77.26825, 77.24647, 77.23536, 77.22976)}


@pytest.mark.parametrize("max_cell", [0.001, 0.0005, 0.00025])
def test_transient_cooling_slab_reproduces_the_references_table(fem, max_cell):
    """A slab at 100 C whose faces go to 0 C (transient heat, heat.rs:684-735: the lumped capacity
    over dt on the diagonal, one solver for every step, each step started from the last solution):
    every entry of the reference's table of grids and time steps to half a unit of its last digit,
    and the case file's own grid and step within its check.toml."""
    from femahoy_amd import grid_case
    case = grid_case.load_case(os.path.join(CASES, "grid-cooling-slab/slab/grid.toml"))
    for dt, printed in zip((0.2, 0.1, 0.05, 0.025), COOLING[max_cell]):
        case.max_cell = max_cell
        case.transient["dt"] = dt
        res = grid_case.solve(case, precision="fp64", tol=1e-12)
        print(max_cell, dt, res.method, res.unknowns, res.probes["centre"])
        assert abs(res.probes["centre"] - printed) <= 0.5e-5, (max_cell, dt, res.probes["centre"], printed)
        if max_cell == 0.0005 and dt == 0.1:
            with open(os.path.join(CASES, "grid-cooling-slab", "check.toml")) as f:
                rows = grid_case.check(res, f.read(), case.name)
            assert rows and all(r[-1] for r in rows)


# grid-bimetal-strip (derivation.md, "Mesh study"): the reference's grid solver on 4 grids of the
# strip heated uniformly by 100 C: unknowns and the top surface's curvatures, to 7 digits.
BIMETAL = {2.0: (1406, 1.122349e-3, 1.132433e-3), 1.0: (4301, 1.158686e-3, 1.157552e-3),
           0.5: (16096, 1.158975e-3, 1.157798e-3), 0.25: (62180, 1.159359e-3, 1.158216e-3)}


@pytest.mark.parametrize("max_cell", [2.0, 1.0, 0.5, 0.25])
def test_bimetal_strip_reproduces_the_references_mesh_study(fem, max_cell):
    """Two bonded layers heated by a uniform 100 C (a static step's `temperature`: each cell's
    thermal force from the same change at every corner, elastic.rs:358-404), free but for uz at the
    centre of the bottom, quarter symmetry: the curvatures of the top surface's quadratic fit
    (elastic.rs:1099-1250) match the reference solver's table to half a unit of its 7th digit,
    the top bulges outward, and the case file's own grid passes its check.toml (Timoshenko's
    curvature 1.163636e-3 within 1 %)."""
    from femahoy_amd import grid_case
    case = grid_case.load_case(os.path.join(CASES, "grid-bimetal-strip/strip/grid.toml"))
    case.max_cell = max_cell
    res = grid_case.solve(case, precision="fp64", tol=1e-12)
    unknowns, kxx, kyy = BIMETAL[max_cell]
    print(max_cell, res.method, res.unknowns, res.probes["top.kxx"], res.probes["top.kyy"])
    assert res.unknowns == unknowns
    assert abs(res.probes["top.kxx"] - kxx) <= 0.5e-9 and abs(res.probes["top.kyy"] - kyy) <= 0.5e-9
    assert res.probes["top.shape"] == "convex"
    if max_cell == 1.0:
        with open(os.path.join(CASES, "grid-bimetal-strip", "check.toml")) as f:
            rows = grid_case.check(res, f.read(), case.name)
        # the surface's curvatures, and the reactions of the face that holds uz (elastic.rs:773-839):
        # the thermal load has no net force, so they are 0 to abs_tol
        assert [r[1] for r in rows] == ["top.kxx", "top.kyy", "centre.fz", "centre.fx"] and all(r[-1] for r in rows)
