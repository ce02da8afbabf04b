"""The allocation-free dense solve (fem_dense_solve_device_noalloc) against the allocating one
(fem_dense_solve_device), through the C ABI: bitwise equal results on two calls in a row, the
second smaller, both within the FP64 bar of numpy's solve; the size limit's message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    k = np.exp(rng.standard_normal(n + 1))
    A = np.diag(k[:-1] + k[1:])
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = -k[i + 1]
    M = rng.standard_normal((n, 4))
    return A + 0.05 * M @ M.T, rng.standard_normal(n)


def test_no_alloc_dense_solve_matches_the_allocating_one(fem):
    for n, seed in ((1000, 1), (137, 2)):  # the second call smaller than the first
        A, b = _spd(n, seed)
        print(f"no-alloc dense solve: n={n} seed={seed}", flush=True)
        x_alloc = fem.dense_solve(A, b, entry="fem_dense_solve_device")
        x_free = fem.dense_solve(A, b, entry="fem_dense_solve_device_noalloc")
        assert np.array_equal(x_free, x_alloc)
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(x_free - ref) <= 1e-10 * np.linalg.norm(ref) * max(1.0, np.linalg.cond(A) / 1e4)


def test_no_alloc_dense_solve_refuses_more_than_its_limit(fem):
    A, b = _spd(1025, 3)
    with pytest.raises(fem.FemError) as e:
        fem.dense_solve(A, b, entry="fem_dense_solve_device_noalloc")
    assert str(e.value) == "the allocation-free dense solve holds at most 1024 unknowns; this matrix has 1025"
