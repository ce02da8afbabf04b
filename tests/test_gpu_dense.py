"""GPU checks of the dense coarse level's building blocks against numpy: the FP64 matrix-core
product (its lane layout) and the LDL^T inverse."""
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
