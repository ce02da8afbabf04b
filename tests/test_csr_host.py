"""The host setup of the assembled boundary (fem_solver_new_csr) against the oracle, bit for bit:
the grid multigrid levels of multigrid.rs:84-139 (coarsen, multigrid.rs:329-452: the kept planes,
the geometric and operator weights, the first-seen coarse numbering; the Galerkin products,
sparse.rs:233-292). The oracle restates the same reference functions in its own code; both must
give identical integer structures (prolongation and coarse matrix patterns, coarse points and
components) and identical FP64 values. CPU only (no device call)."""
import numpy as np
import pytest


def heat_problem(oracle, n, length, seed):
    k = np.exp(np.random.default_rng(seed).standard_normal(int(np.prod(n))))
    return oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0})


def grid_of(pr, n, length):
    """The Grid unknowns the reference's heat caller gives (heat.rs: the free points' plane
    indices; planes i * (length / n))."""
    planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = pr.ijk()[pr.free()]
    return planes, pts, np.zeros(len(pts), np.uint8)


def compare_levels(fem, oracle_solver, ours):
    assert oracle_solver.levels() == len(ours)
    for l, lev in enumerate(ours):
        assert oracle_solver.level_n(l) == lev["n"]
        if l + 1 < len(ours):
            # the coarse matrix of level l + 1 and the prolongation from it
            op, oi, ov = oracle_solver.level_p(l).arrays()
            pp, pi, pv, pc = lev["p"]
            assert oracle_solver.level_p(l).cols == pc
            assert np.array_equal(op.astype(np.uint64), pp)
            assert np.array_equal(oi, pi)
            assert np.array_equal(ov, pv)  # bit for bit
        ap, ai, av = oracle_solver.level_matrix(l).arrays()
        assert np.array_equal(ap.astype(np.uint64), lev["a"][0])
        assert np.array_equal(ai, lev["a"][1])
        assert np.array_equal(av, lev["a"][2])


@pytest.mark.parametrize("n,length", [([24, 18, 14], [1.0, 1.0, 1.0]), ([21, 13, 30], [2.0, 0.7, 1.3])])
def test_scalar_grid_levels_match_the_oracle_bit_for_bit(fem, oracle, n, length):
    pr = heat_problem(oracle, n, length, seed=5)
    a = pr.reduced()
    ptr, idx, val = a.arrays()
    planes, pts, comp = grid_of(pr, n, length)
    ours = fem.csr_levels(ptr, idx, val, fem.Grid(planes, pts, comp))
    ref = oracle.Solver.grid(a, planes, pts, comp, method="multigrid")
    assert len(ours) >= 2
    compare_levels(fem, ref, ours)
    # operator-dependent weights occur (lognormal k): some prolongation values are not 1/2^k
    pv = ours[0]["p"][2]
    assert np.any(np.abs(pv * 64 - np.round(pv * 64)) > 1e-9)


def test_graded_planes_give_the_geometric_weights_of_the_reference(fem, oracle):
    """The same matrix on a graded grid: where the operator weights do not apply (next to the
    fixed planes, where a row has no coupling below or above) the weights follow the planes."""
    n, length = [24, 16, 12], [1.0, 1.0, 1.0]
    pr = heat_problem(oracle, n, length, seed=9)
    a = pr.reduced()
    ptr, idx, val = a.arrays()
    _, pts, comp = grid_of(pr, n, length)
    planes = [np.cumsum(np.r_[0.0, 1.0 + 0.3 * np.sin(np.arange(m))]) for m in n]
    ours = fem.csr_levels(ptr, idx, val, fem.Grid(planes, pts, comp))
    ref = oracle.Solver.grid(a, planes, pts, comp, method="multigrid")
    compare_levels(fem, ref, ours)


def test_vector_grid_levels_keep_the_linear_weights(fem, oracle):
    """A displacement-like matrix (the heat matrix times a 3 x 3 SPD block, unknown 3 i + c): the
    reference keeps the geometric weights for more than one component."""
    import scipy.sparse as sp

    n, length = [16, 12, 10], [1.0, 1.0, 1.0]
    pr = heat_problem(oracle, n, length, seed=3)
    k = pr.reduced().to_scipy().tocsr()
    blk = np.array([[2.0, 0.5, 0.0], [0.5, 2.0, 0.5], [0.0, 0.5, 2.0]])
    m = sp.kron(k, sp.csr_matrix(blk)).tocsr()
    m.sort_indices()
    ptr, idx, val = m.indptr.astype(np.int64), m.indices.astype(np.uint32), m.data.astype(np.float64)
    planes, pts1, _ = grid_of(pr, n, length)
    pts = np.repeat(pts1, 3, axis=0)
    comp = np.tile(np.arange(3, dtype=np.uint8), len(pts1))
    a = oracle.Csr.from_arrays(m.shape[0], ptr, idx, val)
    ours = fem.csr_levels(ptr, idx, val, fem.Grid(planes, pts, comp))
    ref = oracle.Solver.grid(a, planes, pts, comp, method="multigrid")
    assert len(ours) >= 2
    compare_levels(fem, ref, ours)
    for lev in ours[1:]:
        assert set(np.unique(lev["comp"])) == {0, 1, 2}


def test_the_level_loop_stops_as_the_reference_does(fem, oracle):
    """Below 3000 unknowns no coarsening (multigrid.rs:91); max_levels stops earlier; an axis of 2
    planes is kept whole."""
    n, length = [10, 10, 10], [1.0, 1.0, 1.0]
    pr = heat_problem(oracle, n, length, seed=1)
    ptr, idx, val = pr.reduced().arrays()
    planes, pts, comp = grid_of(pr, n, length)
    assert len(fem.csr_levels(ptr, idx, val, fem.Grid(planes, pts, comp))) == 1  # 891 unknowns
    assert len(fem.csr_levels(ptr, idx, val, fem.Grid(planes, pts, comp), coarsest=100)) >= 2
    n2 = [40, 40, 1]
    pr2 = heat_problem(oracle, n2, length, seed=2)
    ptr2, idx2, val2 = pr2.reduced().arrays()
    planes2, pts2, comp2 = grid_of(pr2, n2, length)
    lv = fem.csr_levels(ptr2, idx2, val2, fem.Grid(planes2, pts2, comp2), coarsest=50)
    assert all(set(np.unique(l["points"][:, 2])) <= {0, 1} for l in lv)
    lv2 = fem.csr_levels(ptr2, idx2, val2, fem.Grid(planes2, pts2, comp2), coarsest=50, max_levels=2)
    assert len(lv2) == 2


def test_bad_input_is_refused_with_a_message(fem):
    g = fem.Grid([np.array([0.0, 1.0]), np.array([0.0]), np.array([0.0])], np.array([[0, 0, 0], [1, 0, 0]]),
                 np.zeros(2, np.uint8))
    # row 0 lacks its diagonal entry
    with pytest.raises(fem.FemError, match="diagonal"):
        fem.csr_levels(np.array([0, 1, 3]), np.array([1, 0, 1], np.uint32), np.ones(3), g)
    # the columns of row 0 are not increasing
    with pytest.raises(fem.FemError, match="increasing"):
        fem.csr_levels(np.array([0, 2, 4]), np.array([1, 0, 0, 1], np.uint32), np.ones(4), g)
    # a grid point outside the planes
    bad = fem.Grid(g.planes, np.array([[0, 0, 0], [2, 0, 0]]), g.comp)
    with pytest.raises(fem.FemError, match="outside"):
        fem.csr_levels(np.array([0, 2, 4]), np.array([0, 1, 0, 1], np.uint32), np.array([2.0, -1, -1, 2]), bad)
