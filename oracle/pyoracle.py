"""TEST INFRASTRUCTURE ONLY: ctypes access to the oracle (oracle/liboracle.so).

The oracle is a single-threaded FP64 CPU restatement of the reference's `fem` solve path
(crates/fem/src/{sparse,order,symbolic,cholesky,multigrid,pmultigrid,solve,tet10,hex,heat,prism}.rs).
Only tests/, the smoke check and the CPU-baseline leg of bench.py import this module, and only
as the checker: the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_SZ = C.c_size_t
_I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile the oracle from its sources (g++, -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "orc_csr_pattern": (_P, [_SZ, _SZ, _SZ, _U32P]),
        "orc_csr_from": (_P, [_SZ, _I64P, _U32P, _F64P]),
        "orc_csr_add_element": (C.c_int, [_P, _U32P, _SZ, _F64P, C.c_char_p, _SZ]),
        "orc_csr_n": (_SZ, [_P]),
        "orc_csr_nnz": (_SZ, [_P]),
        "orc_csr_get": (None, [_P, _I64P, _U32P, _F64P]),
        "orc_csr_mul": (None, [_P, _F64P, _F64P]),
        "orc_csr_diagonal": (None, [_P, _F64P]),
        "orc_csr_submatrix": (_P, [_P, _SZ, _I64P]),
        "orc_csr_free": (None, [_P]),
        "orc_rect_from": (_P, [_SZ, _SZ, _I64P, _U32P, _F64P]),
        "orc_rect_rows": (_SZ, [_P]),
        "orc_rect_cols": (_SZ, [_P]),
        "orc_rect_nnz": (_SZ, [_P]),
        "orc_rect_get": (None, [_P, _I64P, _U32P, _F64P]),
        "orc_rect_mul": (None, [_P, _F64P, _F64P]),
        "orc_rect_mul_transpose": (None, [_P, _F64P, _F64P]),
        "orc_rect_galerkin": (_P, [_P, _P]),
        "orc_rect_free": (None, [_P]),
        "orc_nested_dissection": (None, [_P, _F64P, _I64P]),
        "orc_symbolic": (_P, [_P, _I64P]),
        "orc_symbolic_supernodes": (_SZ, [_P]),
        "orc_symbolic_entries": (C.c_uint64, [_P]),
        "orc_symbolic_flops": (C.c_double, [_P]),
        "orc_symbolic_largest_front": (C.c_uint64, [_P]),
        "orc_symbolic_perm": (None, [_P, _I64P]),
        "orc_symbolic_free": (None, [_P]),
        "orc_symbolic_rows_len": (_SZ, [_P]),
        "orc_symbolic_structure": (None, [_P, _I64P, _I64P, np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"),
                                          _I64P]),
        "orc_cholesky": (_P, [_P, _P, C.c_char_p, _SZ]),
        "orc_cholesky_solve": (None, [_P, _F64P]),
        "orc_cholesky_entries": (_SZ, [_P]),
        "orc_cholesky_free": (None, [_P]),
        "orc_conduction": (None, [_F64P, _F64P, _F64P]),
        "orc_tet_gradients": (None, [_F64P, _F64P, _F64P]),
        "orc_tet_conduction": (None, [_F64P, C.c_double, _F64P, _F64P]),
        "orc_orthotropic_stiffness": (C.c_int, [_F64P, _F64P, _F64P, _F64P, _F64P, C.c_char_p, _SZ]),
        "orc_isotropic": (None, [C.c_double, C.c_double, _F64P, _F64P, _F64P]),
        "orc_tet10_stiffness": (None, [_F64P, _F64P, _F64P, _F64P]),
        "orc_tet10_mass": (None, [_F64P, C.c_double, _F64P]),
        "orc_tet4_stiffness": (None, [_F64P, _F64P, _F64P, _F64P]),
        "orc_split_prism": (None, [_U64P, _I64P]),
        "orc_quadratic": (_SZ, [_SZ, _F64P, _SZ, _U32P, _U32P, _U32P]),
        "orc_prolongation": (_P, [_SZ, _F64P, _SZ, _SZ, _U32P, _SZ, _I64P, _F64P]),
        "orc_problem_heat": (_P, [C.c_int, _I32P, _F64P, _F64P, _I32P, _F64P, C.c_double, C.c_char_p, _SZ]),
        "orc_problem_elastic": (_P, [C.c_int, _I32P, _F64P, _F64P, _F64P, _I32P, _F64P, _F64P, C.c_char_p, _SZ]),
        "orc_problem_sizes": (None, [_P, np.ctypeslib.ndpointer(np.uintp, flags="C_CONTIGUOUS")]),
        "orc_problem_points": (None, [_P, _F64P]),
        "orc_problem_ijk": (None, [_P, _U32P]),
        "orc_problem_cells": (None, [_P, _U32P]),
        "orc_problem_edges": (None, [_P, _U32P]),
        "orc_problem_free": (None, [_P, _I64P]),
        "orc_problem_rhs": (None, [_P, _F64P]),
        "orc_problem_load": (None, [_P, _F64P]),
        "orc_problem_fixed": (None, [_P, _U8P, _F64P]),
        "orc_problem_matrix": (_P, [_P]),
        "orc_problem_reduced": (_P, [_P]),
        "orc_problem_prolongation": (_P, [_P]),
        "orc_problem_coarse_xyz": (_SZ, [_P, C.c_void_p]),
        "orc_problem_free_handle": (None, [_P]),
        "orc_solver_new": (_P, [_P, C.c_int, _SZ, C.c_char_p, _SZ]),
        "orc_problem_elastic_hex": (_P, [_I32P, _F64P, _F64P, _F64P, _I32P, _F64P, C.c_char_p, _SZ]),
        "orc_problem_cell_count": (_SZ, [_P]),
        "orc_problem_cell_matrices": (None, [_P, _U32P, _F64P]),
        "orc_hex_stiffness": (None, [_F64P, C.c_double, C.c_double, _F64P, C.c_char_p, _SZ]),
        "orc_solver_new_cells": (_P, [_P, C.c_int, _SZ, C.c_char_p, _SZ]),
        "orc_by_cells_step": (_P, [_P, _U32P, np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"),
                                   C.POINTER(_SZ)]),
        "orc_coarsen": (_P, [_P, _F64P, _SZ, _F64P, _SZ, _F64P, _SZ, _U32P,
                             np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"), _U32P,
                             np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]),
        "orc_grid_solver_new": (_P, [_P, _F64P, _SZ, _F64P, _SZ, _F64P, _SZ, _U32P,
                                     np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"), C.c_int, _SZ,
                                     C.c_char_p, _SZ]),
        "orc_solver_method": (C.c_int, [_P]),
        "orc_solver_levels": (_SZ, [_P]),
        "orc_solver_level_n": (_SZ, [_P, _SZ]),
        "orc_solver_level_p": (_P, [_P, _SZ]),
        "orc_solver_level_matrix": (_P, [_P, _SZ]),
        "orc_solver_precondition": (C.c_int, [_P, _F64P, _F64P]),
        "orc_solver_solve": (C.c_int, [_P, _F64P, _F64P, _SZ, C.c_double, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double), C.c_char_p, _SZ]),
        "orc_solver_free": (None, [_P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class OracleError(RuntimeError):
    """An error of the oracle, with the reference's text."""


def _err() -> C.Array:
    return C.create_string_buffer(1024)


# ------------------------------------------------------------------ Csr / Rect
class Csr:
    """A square CSR (sparse.rs:12-17) owned by the oracle."""

    def __init__(self, handle):
        if not handle:
            raise OracleError("no matrix")
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_csr_free(self.h)
            self.h = None

    @staticmethod
    def pattern(n: int, elements) -> "Csr":
        e = np.ascontiguousarray(elements, dtype=np.uint32)
        return Csr(lib().orc_csr_pattern(n, e.shape[0], e.shape[1], e.ravel()))

    @staticmethod
    def from_arrays(n, ptr, idx, val) -> "Csr":
        return Csr(lib().orc_csr_from(n, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.uint32),
                                      np.ascontiguousarray(val, np.float64)))

    def add_element(self, nodes, ke) -> None:
        nodes = np.ascontiguousarray(nodes, np.uint32)
        err = _err()
        if lib().orc_csr_add_element(self.h, nodes, len(nodes), np.ascontiguousarray(ke, np.float64).ravel(), err, 1024):
            raise OracleError(err.value.decode())

    @property
    def n(self) -> int:
        return lib().orc_csr_n(self.h)

    def arrays(self):
        n, nnz = self.n, lib().orc_csr_nnz(self.h)
        ptr = np.zeros(n + 1, np.int64)
        idx = np.zeros(nnz, np.uint32)
        val = np.zeros(nnz, np.float64)
        lib().orc_csr_get(self.h, ptr, idx, val)
        return ptr, idx, val

    def mul(self, x):
        y = np.zeros(self.n)
        lib().orc_csr_mul(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def diagonal(self):
        d = np.zeros(self.n)
        lib().orc_csr_diagonal(self.h, d)
        return d

    def submatrix(self, keep) -> "Csr":
        keep = np.ascontiguousarray(keep, np.int64)
        return Csr(lib().orc_csr_submatrix(self.h, len(keep), keep))

    def to_scipy(self):
        import scipy.sparse as sp
        ptr, idx, val = self.arrays()
        return sp.csr_matrix((val, idx.astype(np.int64), ptr), shape=(self.n, self.n))


class Rect:
    """A rectangular CSR, the prolongation (sparse.rs:202-208)."""

    def __init__(self, handle):
        if not handle:
            raise OracleError("no prolongation")
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rect_free(self.h)
            self.h = None

    @staticmethod
    def from_arrays(rows, cols, ptr, idx, val) -> "Rect":
        return Rect(lib().orc_rect_from(rows, cols, np.ascontiguousarray(ptr, np.int64),
                                        np.ascontiguousarray(idx, np.uint32), np.ascontiguousarray(val, np.float64)))

    @property
    def rows(self):
        return lib().orc_rect_rows(self.h)

    @property
    def cols(self):
        return lib().orc_rect_cols(self.h)

    def arrays(self):
        nnz = lib().orc_rect_nnz(self.h)
        ptr = np.zeros(self.rows + 1, np.int64)
        idx = np.zeros(nnz, np.uint32)
        val = np.zeros(nnz, np.float64)
        lib().orc_rect_get(self.h, ptr, idx, val)
        return ptr, idx, val

    def galerkin(self, a: Csr) -> Csr:
        return Csr(lib().orc_rect_galerkin(self.h, a.h))

    def mul(self, x):
        y = np.zeros(self.rows)
        lib().orc_rect_mul(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def mul_transpose(self, x):
        y = np.zeros(self.cols)
        lib().orc_rect_mul_transpose(self.h, np.ascontiguousarray(x, np.float64), y)
        return y


def by_cells_step(problem: "Problem"):
    """One coarsen + by_cells step on a hex elasticity problem (multigrid.rs:96-125): the coarse
    matrix (Csr) and the coarse points (n x 3) and components."""
    nf = problem.sizes.n_free
    cp = np.zeros(3 * nf, np.uint32)
    cc = np.zeros(nf, np.uint8)
    cn = C.c_size_t()
    h = lib().orc_by_cells_step(problem.h, cp, cc, C.byref(cn))
    if not h:
        return None
    m = int(cn.value)
    return Csr(h), cp[:3 * m].reshape(-1, 3), cc[:m]


def hex_stiffness(size, e, nu):
    """hex.rs:199-279: the condensed 24 x 24 element of a box cell, isotropic E and nu."""
    out = np.zeros(576)
    err = _err()
    lib().orc_hex_stiffness(np.asarray(size, np.float64), float(e), float(nu), out, err, 1024)
    if err.value:
        raise OracleError(err.value.decode())
    return out.reshape(24, 24)


def coarsen(a: Csr, planes, points, comp):
    """One step of coarsen (multigrid.rs:329-452): the prolongation (Rect) and the coarse points
    and components, or None when no axis has more than 2 planes."""
    pl = [np.ascontiguousarray(p, np.float64) for p in planes]
    cp = np.zeros(3 * a.n, np.uint32)
    cc = np.zeros(a.n, np.uint8)
    h = lib().orc_coarsen(a.h, pl[0], pl[0].size, pl[1], pl[1].size, pl[2], pl[2].size,
                          np.ascontiguousarray(points, np.uint32).reshape(-1), np.ascontiguousarray(comp, np.uint8),
                          cp, cc)
    if not h:
        return None
    r = Rect(h)
    return r, cp[:3 * r.cols].reshape(-1, 3), cc[:r.cols]


def nested_dissection(a: Csr, xyz) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    out = np.zeros(a.n, np.int64)
    lib().orc_nested_dissection(a.h, xyz, out)
    return out


class Symbolic:
    def __init__(self, a: Csr, order):
        self.h = lib().orc_symbolic(a.h, np.ascontiguousarray(order, np.int64))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_symbolic_free(self.h)
            self.h = None

    def entries(self):
        return lib().orc_symbolic_entries(self.h)

    def flops(self):
        return lib().orc_symbolic_flops(self.h)

    def supernodes(self):
        return lib().orc_symbolic_supernodes(self.h)

    def largest_front(self):
        return lib().orc_symbolic_largest_front(self.h)

    def structure(self):
        """(perm, start, rows_ptr, rows, parent) of the supernodal factor (symbolic.rs:41-150);
        parent -1 at a root."""
        ns = self.supernodes()
        n = int(lib().orc_symbolic_rows_len(self.h))
        start, rptr = np.zeros(ns + 1, np.int64), np.zeros(ns + 1, np.int64)
        rows, parent = np.zeros(max(n, 1), np.uint32), np.zeros(max(ns, 1), np.int64)
        lib().orc_symbolic_structure(self.h, start, rptr, rows, parent)
        perm = np.zeros(int(start[-1]), np.int64)
        lib().orc_symbolic_perm(self.h, perm)
        return perm, start, rptr, rows[:n], parent[:ns]


class Cholesky:
    def __init__(self, sym: Symbolic, a: Csr):
        err = _err()
        self.h = lib().orc_cholesky(sym.h, a.h, err, 1024)
        if not self.h:
            raise OracleError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cholesky_free(self.h)
            self.h = None

    def solve(self, b):
        x = np.array(b, dtype=np.float64, copy=True)
        lib().orc_cholesky_solve(self.h, x)
        return x

    def entries(self):
        return lib().orc_cholesky_entries(self.h)


# ------------------------------------------------------------------ elements
def conduction(s, k):
    out = np.zeros(64)
    lib().orc_conduction(np.asarray(s, np.float64), np.asarray(k, np.float64), out)
    return out.reshape(8, 8)


def tet_gradients(p):
    g = np.zeros(12)
    v = np.zeros(1)
    lib().orc_tet_gradients(np.ascontiguousarray(p, np.float64).ravel(), g, v)
    return g.reshape(4, 3), float(v[0])


def tet_conduction(g, v, k):
    out = np.zeros(16)
    lib().orc_tet_conduction(np.ascontiguousarray(g, np.float64).ravel(), v, np.asarray(k, np.float64), out)
    return out.reshape(4, 4)


def orthotropic_stiffness(e, nu, g):
    normal = np.zeros(9)
    shear = np.zeros(3)
    err = _err()
    if lib().orc_orthotropic_stiffness(np.asarray(e, np.float64), np.asarray(nu, np.float64),
                                       np.asarray(g, np.float64), normal, shear, err, 1024):
        raise OracleError(err.value.decode())
    return normal.reshape(3, 3), shear


def isotropic_stiffness(e, nu):
    oe, onu, og = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().orc_isotropic(e, nu, oe, onu, og)
    return orthotropic_stiffness(oe, onu, og)


def tet10_stiffness(p, normal, shear):
    out = np.zeros(900)
    lib().orc_tet10_stiffness(np.ascontiguousarray(p, np.float64).ravel(), np.ascontiguousarray(normal, np.float64).ravel(),
                              np.asarray(shear, np.float64), out)
    return out.reshape(30, 30)


def tet10_mass(p, density):
    out = np.zeros(900)
    lib().orc_tet10_mass(np.ascontiguousarray(p, np.float64).ravel(), density, out)
    return out.reshape(30, 30)


def tet4_stiffness(p, normal, shear):
    out = np.zeros(144)
    lib().orc_tet4_stiffness(np.ascontiguousarray(p, np.float64).ravel(), np.ascontiguousarray(normal, np.float64).ravel(),
                             np.asarray(shear, np.float64), out)
    return out.reshape(12, 12)


def split_prism(g):
    out = np.zeros(12, np.int64)
    lib().orc_split_prism(np.asarray(g, np.uint64), out)
    return out.reshape(3, 4)


def quadratic(points, tets):
    points = np.ascontiguousarray(points, np.float64)
    tets = np.ascontiguousarray(tets, np.uint32)
    nt = tets.shape[0]
    edges = np.zeros(2 * 6 * nt, np.uint32)
    t10 = np.zeros(10 * nt, np.uint32)
    ne = lib().orc_quadratic(points.shape[0], points.ravel(), nt, tets.ravel(), edges, t10)
    return edges[: 2 * ne].reshape(ne, 2), t10.reshape(nt, 10)


def prolongation(points, corners, edges, free):
    points = np.ascontiguousarray(points, np.float64)
    edges = np.ascontiguousarray(edges, np.uint32)
    free = np.ascontiguousarray(free, np.int64)
    xyz = np.zeros(3 * 3 * corners)
    h = lib().orc_prolongation(points.shape[0], points.ravel(), corners, edges.shape[0], edges.ravel(), len(free), free, xyz)
    p = Rect(h)
    return p, xyz[: 3 * p.cols].reshape(-1, 3)


# ------------------------------------------------------------------ structured box problems
FACES = ("x0", "x1", "y0", "y1", "z0", "z1")


@dataclass
class ProblemArrays:
    n_points: int
    comps: int
    n_unknowns: int
    n_free: int
    npe: int
    n_cells: int
    corners: int
    n_edges: int


class Problem:
    """A structured box model built the way the reference callers build theirs."""

    def __init__(self, handle):
        if not handle:
            raise OracleError("no problem")
        self.h = handle
        s = np.zeros(8, np.uintp)
        lib().orc_problem_sizes(self.h, s)
        self.sizes = ProblemArrays(*[int(v) for v in s])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_free_handle(self.h)
            self.h = None

    @staticmethod
    def heat(n, length, k, fixed=None, value=None, source=0.0, tet4=False) -> "Problem":
        fixed_a = np.zeros(6, np.int32)
        value_a = np.zeros(6, np.float64)
        for f, v in (fixed or {}).items():
            fixed_a[FACES.index(f)] = 1
            value_a[FACES.index(f)] = v
        err = _err()
        h = lib().orc_problem_heat(1 if tet4 else 0, np.asarray(n, np.int32), np.asarray(length, np.float64),
                                   np.ascontiguousarray(k, np.float64), fixed_a, value_a, source, err, 1024)
        if not h:
            raise OracleError(err.value.decode())
        return Problem(h)

    @staticmethod
    def elastic(order, n, length, e, nu, clamped=(), body=(0, 0, 0), traction=None) -> "Problem":
        nt = 6 * int(np.prod(n))
        e = np.broadcast_to(np.asarray(e, np.float64), (nt,)).copy()
        nu = np.broadcast_to(np.asarray(nu, np.float64), (nt,)).copy()
        cl = np.zeros(6, np.int32)
        for f in clamped:
            cl[FACES.index(f)] = 1
        tr = np.zeros(18)
        for f, t in (traction or {}).items():
            tr[3 * FACES.index(f): 3 * FACES.index(f) + 3] = t
        err = _err()
        h = lib().orc_problem_elastic(order, np.asarray(n, np.int32), np.asarray(length, np.float64), e, nu, cl,
                                      np.asarray(body, np.float64), tr, err, 1024)
        if not h:
            raise OracleError(err.value.decode())
        return Problem(h)

    @staticmethod
    def elastic_hex(n, length, e, nu, clamped=(), body=(0, 0, 0)) -> "Problem":
        """Box-grid linear elasticity on hex8 cells (elastic.rs:287-563), E and nu per cell."""
        nc = int(np.prod(n))
        e = np.broadcast_to(np.asarray(e, np.float64), (nc,)).copy()
        nu = np.broadcast_to(np.asarray(nu, np.float64), (nc,)).copy()
        cl = np.zeros(6, np.int32)
        for f in clamped:
            cl[FACES.index(f)] = 1
        err = _err()
        h = lib().orc_problem_elastic_hex(np.asarray(n, np.int32), np.asarray(length, np.float64), e, nu, cl,
                                          np.asarray(body, np.float64), err, 1024)
        if not h:
            raise OracleError(err.value.decode())
        return Problem(h)

    def cell_matrices(self):
        """The cells of a hex elasticity problem (multigrid.rs:67-70): origins (count x 3) and
        element matrices (count x 24 x 24)."""
        c = lib().orc_problem_cell_count(self.h)
        origin = np.zeros(3 * c, np.uint32)
        k = np.zeros(576 * c)
        lib().orc_problem_cell_matrices(self.h, origin, k)
        return origin.reshape(c, 3), k.reshape(c, 24, 24)

    def points(self):
        out = np.zeros(3 * self.sizes.n_points)
        lib().orc_problem_points(self.h, out)
        return out.reshape(-1, 3)

    def ijk(self):
        out = np.zeros(3 * self.sizes.n_points, np.uint32)
        lib().orc_problem_ijk(self.h, out)
        return out.reshape(-1, 3)

    def cells(self):
        out = np.zeros(self.sizes.n_cells * self.sizes.npe, np.uint32)
        lib().orc_problem_cells(self.h, out)
        return out.reshape(self.sizes.n_cells, self.sizes.npe)

    def edges(self):
        out = np.zeros(2 * self.sizes.n_edges, np.uint32)
        lib().orc_problem_edges(self.h, out)
        return out.reshape(-1, 2)

    def free(self):
        out = np.zeros(self.sizes.n_free, np.int64)
        lib().orc_problem_free(self.h, out)
        return out

    def rhs(self):
        out = np.zeros(self.sizes.n_free)
        lib().orc_problem_rhs(self.h, out)
        return out

    def load(self):
        out = np.zeros(self.sizes.n_unknowns)
        lib().orc_problem_load(self.h, out)
        return out

    def fixed(self):
        isf = np.zeros(self.sizes.n_unknowns, np.uint8)
        val = np.zeros(self.sizes.n_unknowns)
        lib().orc_problem_fixed(self.h, isf, val)
        return isf, val

    def matrix(self) -> Csr:
        return Csr(lib().orc_problem_matrix(self.h))

    def reduced(self) -> Csr:
        return Csr(lib().orc_problem_reduced(self.h))

    def prolongation(self) -> Rect:
        return Rect(lib().orc_problem_prolongation(self.h))


class Solver:
    """solve.rs Solver on a problem's reduced matrix and unknowns."""

    def __init__(self, problem: Problem, method=None, crossover=3000):
        m = {None: -1, "direct": 0, "multigrid": 1}[method]
        err = _err()
        self.h = lib().orc_solver_new(problem.h, m, crossover, err, 1024)
        if not self.h:
            raise OracleError(err.value.decode())
        self.n = problem.sizes.n_free

    @classmethod
    def with_cells(cls, problem: Problem, method=None, crossover=3000) -> "Solver":
        """Solver::new(.., Some(cells)) on a hex elasticity problem: the coarse matrices of
        by_cells (multigrid.rs:458-600)."""
        m = {None: -1, "direct": 0, "multigrid": 1}[method]
        err = _err()
        self = cls.__new__(cls)
        self.h = lib().orc_solver_new_cells(problem.h, m, crossover, err, 1024)
        if not self.h:
            raise OracleError(err.value.decode())
        self.n = problem.sizes.n_free
        return self

    @classmethod
    def grid(cls, a: "Csr", planes, points, comp, method=None, crossover=3000) -> "Solver":
        """solve.rs Solver::new(a, Unknowns::Grid(level), ...) on an assembled matrix."""
        m = {None: -1, "direct": 0, "multigrid": 1}[method]
        pl = [np.ascontiguousarray(p, np.float64) for p in planes]
        err = _err()
        self = cls.__new__(cls)
        self.h = lib().orc_grid_solver_new(a.h, pl[0], pl[0].size, pl[1], pl[1].size, pl[2], pl[2].size,
                                           np.ascontiguousarray(points, np.uint32).reshape(-1),
                                           np.ascontiguousarray(comp, np.uint8), m, crossover, err, 1024)
        if not self.h:
            raise OracleError(err.value.decode())
        self.n = a.n
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_free(self.h)
            self.h = None

    @property
    def method(self) -> str:
        return ("direct", "multigrid")[lib().orc_solver_method(self.h)]

    def levels(self) -> int:
        return lib().orc_solver_levels(self.h)

    def level_n(self, l) -> int:
        return lib().orc_solver_level_n(self.h, l)

    def level_p(self, l) -> Rect:
        return Rect(lib().orc_solver_level_p(self.h, l))

    def level_matrix(self, l) -> Csr:
        return Csr(lib().orc_solver_level_matrix(self.h, l))

    def precondition(self, r):
        z = np.zeros(self.n)
        lib().orc_solver_precondition(self.h, np.ascontiguousarray(r, np.float64), z)
        return z

    def solve(self, b, x0=None, max_iter=500, tol=1e-10):
        x = np.zeros(self.n) if x0 is None else np.array(x0, np.float64, copy=True)
        it = C.c_uint64()
        res = C.c_double()
        err = _err()
        if lib().orc_solver_solve(self.h, np.ascontiguousarray(b, np.float64), x, max_iter, tol, C.byref(it),
                                  C.byref(res), err, 1024):
            raise OracleError(err.value.decode())
        return x, int(it.value), float(res.value)
