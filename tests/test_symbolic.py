"""The library's host nested dissection and symbolic counts (the Crossover::Flops analysis of
solve.rs:106-118 and 158-170) against the reference's known answer (order.rs:186), closed forms, a
dense factor's nonzeros, and the oracle's order and counts bit for bit. CPU only."""
import numpy as np
import pytest


def _pattern_arrays(pairs, n):
    rows = [set([i]) for i in range(n)]
    for a, b in pairs:
        rows[a].add(b)
        rows[b].add(a)
    ptr = np.zeros(n + 1, np.uint64)
    idx = []
    for i, r in enumerate(rows):
        idx += sorted(r)
        ptr[i + 1] = len(idx)
    return ptr, np.array(idx, np.uint32)


def test_the_line_of_200_nodes_has_its_middle_last(fem):
    """order.rs:186-199: a line of 200 nodes; the order is a permutation and the first split's
    separator (node 99) comes last."""
    n = 200
    ptr, idx = _pattern_arrays([(i, i + 1) for i in range(n - 1)], n)
    xyz = np.array([[i, 0.0, 0.0] for i in range(n)])
    order, entries, flops, front, sn = fem.symbolic(ptr, idx, xyz)
    assert sorted(order.tolist()) == list(range(n))
    assert order[-1] == 99


@pytest.mark.parametrize("n", [10, 64])
def test_a_line_in_its_own_order_has_the_tridiagonal_counts(fem, n):
    """At most 64 nodes are not split (order.rs:15), so the order is the natural one; L of a
    tridiagonal matrix has 2 entries per column but the last: entries 2n - 1, flops 4(n - 1) + 1,
    and n - 1 supernodes (the last two columns share one), largest front 2^2."""
    ptr, idx = _pattern_arrays([(i, i + 1) for i in range(n - 1)], n)
    xyz = np.array([[i, 0.0, 0.0] for i in range(n)])
    order, entries, flops, front, sn = fem.symbolic(ptr, idx, xyz)
    assert order.tolist() == list(range(n))
    assert entries == 2 * n - 1
    assert flops == 4 * (n - 1) + 1
    assert sn == n - 1
    assert front == 4


def test_entries_are_the_nonzeros_of_a_dense_factor(fem):
    """A 9 x 7 x 5 grid Laplacian (7-point) with random positive couplings: the entries of the
    symbolic factor equal the nonzeros of numpy's Cholesky factor of P A P^T in the same order."""
    nx, ny, nz = 9, 7, 5
    n = nx * ny * nz
    node = lambda i, j, k: i + nx * (j + ny * k)
    pairs = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if i + 1 < nx:
                    pairs.append((node(i, j, k), node(i + 1, j, k)))
                if j + 1 < ny:
                    pairs.append((node(i, j, k), node(i, j + 1, k)))
                if k + 1 < nz:
                    pairs.append((node(i, j, k), node(i, j, k + 1)))
    ptr, idx = _pattern_arrays(pairs, n)
    xyz = np.array([[i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)], np.float64)
    order, entries, flops, front, sn = fem.symbolic(ptr, idx, xyz)
    rng = np.random.default_rng(3)
    A = np.zeros((n, n))
    for a, b in pairs:
        w = rng.uniform(0.5, 2.0)
        A[a, b] = A[b, a] = -w
        A[a, a] += w
        A[b, b] += w
    A += 0.1 * np.eye(n)
    L = np.linalg.cholesky(A[np.ix_(order, order)])
    assert entries == int(np.count_nonzero(np.abs(L) > 1e-300))


def _heat(oracle, n, length, seed):
    k = np.exp(np.random.default_rng(seed).standard_normal(int(np.prod(n))))
    pr = oracle.Problem.heat(n, length, k, fixed={"x0": 0.0, "x1": 1.0})
    planes = [(length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = pr.ijk()[pr.free()]
    xyz = np.stack([planes[d][pts[:, d]] for d in range(3)], axis=1)
    return pr.reduced(), xyz


@pytest.mark.parametrize("n,length", [([12, 10, 8], [1.0, 1.0, 1.0]), ([30, 6, 2], [3.0, 0.6, 0.1]),
                                      ([7, 7, 7], [1.0, 2.0, 3.0])])
def test_order_and_counts_equal_the_oracle_on_heat_grids(fem, oracle, n, length):
    a, xyz = _heat(oracle, n, length, 4)
    ptr, idx, _ = a.arrays()
    order, entries, flops, front, sn = fem.symbolic(ptr.astype(np.uint64), idx, xyz)
    ref = oracle.nested_dissection(a, xyz)
    assert np.array_equal(order, ref)
    s = oracle.Symbolic(a, ref)
    assert (entries, flops, front, sn) == (s.entries(), s.flops(), s.largest_front(), s.supernodes())


@pytest.mark.parametrize("n", [[4, 3, 3], [10, 10, 1]])
def test_order_and_counts_equal_the_oracle_on_tet10(fem, oracle, n):
    """Quadratic unknowns with the coordinates of each unknown's node (elastic_tet.rs:299-300);
    10 x 10 x 1 is a flat plate."""
    nt = 6 * int(np.prod(n))
    pr = oracle.Problem.elastic(10, n, [float(n[0]), float(n[1]), 0.25 * n[2]], np.ones(nt), np.full(nt, 0.3),
                                clamped=["x0"])
    a = pr.reduced()
    xyz = pr.points()[pr.free() // 3]
    ptr, idx, _ = a.arrays()
    order, entries, flops, front, sn = fem.symbolic(ptr.astype(np.uint64), idx, xyz)
    ref = oracle.nested_dissection(a, xyz)
    assert np.array_equal(order, ref)
    s = oracle.Symbolic(a, ref)
    assert (entries, flops, front, sn) == (s.entries(), s.flops(), s.largest_front(), s.supernodes())


def test_grid_indices_give_the_order_of_the_coordinates(fem, oracle):
    """The box solver's crossover orders by grid indices: the order reads coordinates only through
    comparisons along each axis, so any increasing map of each axis gives the same order."""
    a, xyz = _heat(oracle, [11, 9, 6], [1.3, 0.9, 0.4], 5)
    ptr, idx, _ = a.arrays()
    o1 = fem.symbolic(ptr.astype(np.uint64), idx, xyz)[0]
    warped = np.stack([np.round(xyz[:, 0] * 11 / 1.3), np.exp(xyz[:, 1]), xyz[:, 2] ** 3], axis=1)
    o2 = fem.symbolic(ptr.astype(np.uint64), idx, warped)[0]
    assert np.array_equal(o1, o2)


@pytest.mark.parametrize("case", ["heat", "elastic_hex", "tet10"])
def test_the_multifrontal_structure_is_the_oracles(fem, oracle, case):
    """The supernodes of the multifrontal factor and their rows (symbolic.rs:41-150: the postordered
    nested-dissection order, the fundamental supernodes, each one's own columns then its rows below
    them, the supernode tree) equal the oracle's, integer for integer."""
    rng = np.random.default_rng(51)
    if case == "heat":
        n = [9, 7, 6]
        pr = oracle.Problem.heat(n, [1.0, 0.8, 0.6], np.exp(rng.standard_normal(int(np.prod(n)))),
                                 fixed={"x0": 0.0, "x1": 1.0})
    elif case == "elastic_hex":
        n = [8, 4, 3]
        nc = int(np.prod(n))
        pr = oracle.Problem.elastic_hex(n, [4.0, 2.0, 1.5], rng.uniform(1, 2, nc), rng.uniform(0.2, 0.4, nc),
                                        clamped=["x0"])
    else:
        n = [4, 3, 2]
        nt = 6 * int(np.prod(n))
        pr = oracle.Problem.elastic(10, n, [4.0, 3.0, 2.0], rng.uniform(1, 2, nt), rng.uniform(0.2, 0.4, nt),
                                    clamped=["x0"])
    a = pr.reduced()
    ptr, idx, _ = a.arrays()
    comps = 1 if case == "heat" else 3
    xyz = pr.points()[pr.free() // comps]
    perm, start, rptr, rows, sparent = fem.supernodal(ptr, idx, xyz)
    operm, ostart, orptr, orows, oparent = oracle.Symbolic(a, oracle.nested_dissection(a, xyz)).structure()
    assert np.array_equal(perm, operm)
    assert np.array_equal(start, ostart)
    assert np.array_equal(rptr, orptr)
    assert np.array_equal(rows, orows.astype(np.int64))
    assert np.array_equal(sparent, oparent)
    # each supernode's rows start with its own columns and increase
    for s in range(len(start) - 1):
        r = rows[rptr[s]:rptr[s + 1]]
        k = start[s + 1] - start[s]
        assert np.array_equal(r[:k], np.arange(start[s], start[s + 1])) and np.all(np.diff(r) > 0)


@pytest.mark.parametrize("join", [16, 48])
def test_joined_supernodes_keep_the_factor(fem, oracle, join):
    """The multifrontal factor's joins (a supernode that is its parent's last child joins it when the
    two hold at most `join` pivots): the order is the fundamental one; pivot ranges stay contiguous
    and cover every column once; each joined supernode's rows are its pivots then the rows of the
    fundamental supernode it ends in; every fundamental supernode's rows lie in the rows of the
    joined supernode that holds it, and every child's rows past its pivots lie in its parent's rows
    (the extend-add of the factor); join 0 gives the fundamental supernodes unchanged."""
    a, xyz = _heat(oracle, [14, 9, 7], [1.0, 0.8, 0.6], 12)
    ptr, idx, _ = a.arrays()
    ptr = ptr.astype(np.uint64)
    f = fem.supernodal(ptr, idx, xyz)
    z = fem.supernodal(ptr, idx, xyz, join=0)
    assert all(np.array_equal(u, v) for u, v in zip(f, z))
    perm, start, rptr, rows, sparent = fem.supernodal(ptr, idx, xyz, join=join)
    fperm, fstart, frptr, frows, fsparent = f
    assert np.array_equal(perm, fperm)
    ns, fns = len(start) - 1, len(fstart) - 1
    assert ns < fns and start[0] == 0 and start[-1] == len(perm) and np.all(np.diff(start) > 0)
    # every joined supernode starts at a fundamental one's first column and ends at one's last
    assert set(start.tolist()) <= set(fstart.tolist())
    owner = np.searchsorted(start, fstart[:-1], side="right") - 1  # the joined supernode of each fundamental one
    for q in range(ns):
        own = rows[rptr[q]:rptr[q + 1]]
        k = start[q + 1] - start[q]
        assert np.array_equal(own[:k], np.arange(start[q], start[q + 1]))
        assert np.all(np.diff(own) > 0)
        # its last fundamental supernode's rows past its pivots follow
        last = np.searchsorted(fstart, start[q + 1]) - 1
        fk = fstart[last + 1] - fstart[last]
        assert np.array_equal(own[k:], frows[frptr[last] + fk:frptr[last + 1]])
        if k > 1 and len(np.flatnonzero(owner == q)) > 1:
            assert k <= join
        if sparent[q] >= 0:
            p = sparent[q]
            assert p > q
            prow = set(rows[rptr[p]:rptr[p + 1]].tolist())
            assert set(own[k:].tolist()) <= prow
    for s in range(fns):
        q = owner[s]
        assert set(frows[frptr[s]:frptr[s + 1]].tolist()) <= set(rows[rptr[q]:rptr[q + 1]].tolist())
