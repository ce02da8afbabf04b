"""GPU checks of the dense coarse level's building blocks against numpy: the FP64 matrix-core
product (its lane layout), the LDL^T inverse, and the FP32 V-cycle's coarse solve."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hooks(fem):
    L = fem.lib()
    F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.fem_test_gemm_nt.restype = C.c_int
    L.fem_test_gemm_nt.argtypes = [C.c_int, C.c_int, C.c_int, F64P, F64P, F64P, C.c_char_p, C.c_size_t]
    L.fem_test_dense_inverse.restype = C.c_int
    L.fem_test_dense_inverse.argtypes = [C.c_int, F64P, F64P, C.c_char_p, C.c_size_t]
    F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fem_test_dense_solve32.restype = C.c_int
    L.fem_test_dense_solve32.argtypes = [C.c_int, F64P, F32P, F32P, C.c_char_p, C.c_size_t]
    return L


@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (128, 64, 64), (64, 192, 128)])
def test_mfma_gemm_nt_matches_numpy(fem, m, n, k):
    L = _hooks(fem)
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    out = np.zeros((m, n))
    err = C.create_string_buffer(512)
    assert L.fem_test_gemm_nt(m, n, k, A.ravel(), B.ravel(), out.ravel(), err, 512) == 0, err.value
    ref = A @ B.T
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max() * k


@pytest.mark.parametrize("n", [5, 64, 100, 300])
def test_dense_inverse_matches_numpy(fem, n):
    L = _hooks(fem)
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    out = np.zeros((n, n))
    err = C.create_string_buffer(512)
    assert L.fem_test_dense_inverse(n, np.ascontiguousarray(A).ravel(), out.ravel(), err, 512) == 0, err.value
    ref = np.linalg.inv(A)
    assert np.abs(out - ref).max() < 1e-10 * np.abs(ref).max()


def test_dense_inverse_refuses_a_singular_matrix(fem):
    L = _hooks(fem)
    n = 70
    A = np.zeros((n, n))
    for i in range(n - 1):  # a path Laplacian with no shift: singular
        A[i, i] += 1
        A[i + 1, i + 1] += 1
        A[i, i + 1] -= 1
        A[i + 1, i] -= 1
    out = np.zeros((n, n))
    err = C.create_string_buffer(512)
    assert L.fem_test_dense_inverse(n, A.ravel(), out.ravel(), err, 512) != 0
    assert b"not positive definite" in err.value


@pytest.mark.parametrize("n", [5, 100, 517, 1500])
def test_the_fp32_coarse_solve_matches_numpy(fem, n):
    """A graded 1D Laplacian plus a random SPD part (entries of the inverse over several decades):
    the solve from the FP32 inverse is within the FP32 rounding of the entries and of the sums."""
    L = _hooks(fem)
    rng = np.random.default_rng(n)
    k = np.exp(rng.standard_normal(n + 1))
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = k[i] + k[i + 1]
        if i + 1 < n:
            A[i, i + 1] = A[i + 1, i] = -k[i + 1]
    M = rng.standard_normal((n, 8))
    A += 0.01 * M @ M.T
    b = rng.standard_normal(n).astype(np.float32)
    x = np.zeros(n, np.float32)
    err = C.create_string_buffer(512)
    assert L.fem_test_dense_solve32(n, np.ascontiguousarray(A).ravel(), b, x, err, 512) == 0, err.value
    inv = np.linalg.inv(A)
    ref = inv @ b.astype(np.float64)
    # the FP32 rounding of each entry and FP32 sums of up to n / 64 + 6 terms per lane and wave
    bound = (n / 64 + 8) * 2.0 ** -24 * (np.abs(inv) @ np.abs(b))
    assert np.all(np.abs(x - ref) <= bound)


@pytest.mark.parametrize("n", [1, 63, 64, 200, 1111])
def test_dense_solve_device_matches_numpy(fem, n):
    """fem_dense_solve_device (factor, inverse, one refinement step) against numpy's solve: an
    SPD matrix with entries over several decades, error within a few FP64 roundings of cond(A)."""
    rng = np.random.default_rng(100 + n)
    k = np.exp(rng.standard_normal(n + 1))
    A = np.diag(k[:-1] + k[1:])
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = -k[i + 1]
    M = rng.standard_normal((n, 4))
    A += 0.05 * M @ M.T
    b = rng.standard_normal(n)
    x = fem.dense_solve(A, b)
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref) * max(1.0, np.linalg.cond(A) / 1e4)
    assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b) * np.linalg.cond(A)


def test_dense_solve_device_refuses_a_singular_matrix(fem, oracle):
    """The reference's text, with the diagonal in Rust's {:e} (cholesky.rs:94-98): a path
    Laplacian with no shift fails at its last pivot, as the oracle's multifrontal factor does."""
    n = 70
    A = np.zeros((n, n))
    for i in range(n - 1):
        A[i, i] += 1
        A[i + 1, i + 1] += 1
        A[i, i + 1] -= 1
        A[i + 1, i] -= 1
    with pytest.raises(fem.FemError) as e:
        fem.dense_solve(A, np.ones(n))
    text = ("the matrix is not positive definite: the pivot of column 69 is 0 or less to the "
            "precision of the arithmetic (its diagonal is 1e0)")
    assert str(e.value) == text
    # the oracle's multifrontal factor in the natural order stops at the same column, same text
    ptr = np.zeros(n + 1, np.int64)
    idx, val = [], []
    for i in range(n):
        cols = np.nonzero(A[i])[0]
        idx += list(cols)
        val += list(A[i, cols])
        ptr[i + 1] = len(idx)
    a = oracle.Csr.from_arrays(n, ptr, np.array(idx, np.uint32), np.array(val))
    with pytest.raises(oracle.OracleError) as eo:
        oracle.Cholesky(oracle.Symbolic(a, np.arange(n, dtype=np.int64)), a)
    assert str(eo.value) == text
