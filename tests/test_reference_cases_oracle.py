"""The oracle against values the reference's own solver printed (CPU).

The derivation notes of the reference's validation cases tabulate its grid solver's results on
each mesh of a mesh study: validation/cases/grid-heated-bar/derivation.md (the centre temperature,
7 significant digits) and validation/cases/grid-cantilever/derivation.md (the tip deflection, 6).
The oracle's restatement of the same meshes, elements, numbering and loads, solved by its
multifrontal Cholesky (and its line Gauss-Seidel multigrid above 3,000 unknowns, as solve.rs
chooses), must give them to half a unit of the last printed digit. The device solves are checked
against the same values in test_gpu_reference_cases.py.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("max_cell,printed", [(0.00125, 14.779861), (0.000625, 14.745623)])
def test_oracle_heated_bar_centre_temperature_is_the_references(oracle, max_cell, printed):
    n = [round(0.0025 / max_cell), round(0.01 / max_cell), round(0.01 / max_cell)]
    k = np.full(int(np.prod(n)), 20.0)
    pr = oracle.Problem.heat(n, [0.0025, 0.01, 0.01], k, fixed={"y0": 0.0, "z0": 0.0}, source=1e7)
    x, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs())
    ijk = pr.ijk()
    free = pr.free()
    at = [i for i, f in enumerate(free) if tuple(ijk[f]) == (n[0] // 2, n[1], n[2])]
    assert len(at) == 1
    assert abs(x[at[0]] - printed) <= 0.5e-6


@pytest.mark.parametrize("max_cell,unknowns,printed", [(5.0, 540, -0.189395), (2.5, 3000, -0.190113)])
def test_oracle_hex_cantilever_tip_deflection_is_the_references(oracle, max_cell, unknowns, printed):
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
    x, _, _ = oracle.Solver(pr, "direct").solve(pr.rhs() + force[free])
    tip = 3 * node_of[(n[0], n[1] // 2, n[2] // 2)] + 2
    assert abs(x[np.searchsorted(free, tip)] - printed) <= 0.5e-6
