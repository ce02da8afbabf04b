"""femahoy_amd: the iterative solve of femahoy's `fem` crate on one AMD MI355X.

The host side of the drop-in boundary, over the C ABI of include/fem_solver.h
(libfemahoy_amd.so, built in-tree by `make -C femahoy_amd`). It mirrors the reference's
solver interface (crates/fem/src/solve.rs):

    Solver.new(box, method, ...)      Solver::new (solve.rs:92)
    solver.with_max_iter(n)           Solver::with_max_iter (solve.rs:187)
    solver.method()                   Solver::method (solve.rs:191)
    solver.solve(b, x, cancelled)     Solver::solve (solve.rs:199), returning Stats (solve.rs:81)

with the same error texts ("conjugate gradients did not converge: ...", "cancelled"). The
library is the only compute path: if it is missing, importing works but every solver call
raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemahoy_amd.so")

HEAT_HEX8, HEAT_TET4, ELASTIC_TET4, ELASTIC_TET10, ELASTIC_HEX8 = 0, 1, 2, 3, 4
KINDS = {"heat_hex8": HEAT_HEX8, "heat_tet4": HEAT_TET4, "elastic_tet4": ELASTIC_TET4, "elastic_tet10": ELASTIC_TET10,
         "elastic_hex8": ELASTIC_HEX8}
HEX_KINDS = ("heat_hex8", "elastic_hex8")  # one material value per cell (the others: per tetrahedron)
MULTIGRID_ASPECT = 2.0  # elastic.rs:276: cells flatter than this are 'flat' to the FLOPS crossover
METHODS = {None: -1, "direct": 0, "multigrid": 1}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}
FACES = ("x0", "x1", "y0", "y1", "z0", "z1")
# The chain axis order of each Kuhn type, as the device code and fem_model_tet_cells number them.
KUHN_AXES = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))

PCG_TOL = 1e-10        # solve.rs:42
PCG_MAX_ITER = 500     # solve.rs:43
STEADY_DIRECT_LIMIT = 3000  # solve.rs:18


class FemError(RuntimeError):
    """An error of the library, with the reference's text where the reference has one."""


class _Box(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n", C.c_int32 * 3),
        ("length", C.c_double * 3),
        ("coef0", C.POINTER(C.c_double)),
        ("coef1", C.POINTER(C.c_double)),
        ("fixed", C.c_int32 * 6),
        ("value", C.c_double * 6),
        ("source", C.c_double),
        ("body", C.c_double * 3),
        ("traction", C.c_double * 18),
    ]


class _Options(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_iter", C.c_uint64),
        ("method", C.c_int32),
        ("precision", C.c_int32),
        ("crossover", C.c_uint64),
        ("levels", C.c_int32),
        ("cheb_degree", C.c_int32),
        ("cheb_lower", C.c_double),
        ("cheb_upper", C.c_double),
        ("power_steps", C.c_int32),
        ("coarse_max", C.c_uint64),
        ("refine", C.c_int32),
        ("device", C.c_int32),
        ("coarse_galerkin", C.c_int32),
        ("crossover_kind", C.c_int32),
        ("crossover_flops", C.c_double),
        ("crossover_entries", C.c_uint64),
        ("crossover_flat", C.c_int32),
        ("direct_solver", C.c_int32),
    ]


DIRECT_SOLVERS = {"auto": 0, "dense": 1, "multifrontal": 2}


class _Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint64),
        ("residual", C.c_double),
        ("true_residual", C.c_double),
        ("seconds", C.c_double),
        ("setup_seconds", C.c_double),
        ("levels", C.c_uint64),
        ("coarse_unknowns", C.c_uint64),
        ("restarts", C.c_uint64),
        ("method", C.c_int32),
        ("precision", C.c_int32),
        ("hbm_bytes", C.c_uint64),
    ]


CANCEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)

UNKNOWNS = {"grid": 0, "points": 1, "quadratic": 2}


class _Csr(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("cols", C.c_uint64), ("ptr", C.POINTER(C.c_uint64)),
                ("idx", C.POINTER(C.c_uint32)), ("val", C.POINTER(C.c_double))]


class _GridLevel(C.Structure):
    _fields_ = [("planes", C.POINTER(C.c_double) * 3), ("nplanes", C.c_uint64 * 3),
                ("points", C.POINTER(C.c_uint32)), ("comp", C.POINTER(C.c_uint8))]


class _Cells(C.Structure):
    _fields_ = [("count", C.c_uint64), ("origin", C.c_void_p), ("element", C.c_void_p)]


class _Unknowns(C.Structure):
    _fields_ = [("kind", C.c_int32), ("grid", _GridLevel), ("p", _Csr), ("coarse_has_grid", C.c_int32),
                ("coarse_grid", _GridLevel), ("xyz", C.c_void_p), ("cells", C.POINTER(_Cells))]

class _Mesh(C.Structure):
    _fields_ = [("kind", C.c_int32), ("points", C.c_uint64), ("xyz", C.c_void_p), ("cells", C.c_uint64),
                ("conn", C.c_void_p), ("coef0", C.c_void_p), ("coef1", C.c_void_p), ("source", C.c_double),
                ("body", C.c_double * 3), ("fixed", C.c_void_p), ("value", C.c_void_p), ("load", C.c_void_p),
                ("ortho", C.c_void_p)]


class AssemblyStats(C.Structure):
    """fem_assembly_stats: the sizes of an assembled system and the time of its two parts."""

    _fields_ = [("free", C.c_uint64), ("entries", C.c_uint64), ("full_entries", C.c_uint64),
                ("host_seconds", C.c_double), ("device_seconds", C.c_double)]


class _ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_active", C.c_int32), ("device", C.c_int32),
                ("setup_threads", C.c_int32), ("unbatched", C.c_int32)]


class RequestInfo(C.Structure):
    """fem_request_info: the timeline (seconds since the server started) and footprint of a solve."""

    _fields_ = [("submit_time", C.c_double), ("admit_time", C.c_double), ("start_time", C.c_double),
                ("finish_time", C.c_double), ("footprint", C.c_uint64), ("device_bytes", C.c_int64)]


class Profile(C.Structure):
    """fem_profile: device time of the hot launches of the last solve (measurement only)."""

    _fields_ = [("smoother_launches", C.c_uint64), ("smoother_seconds", C.c_double), ("smoother_bytes", C.c_uint64),
                ("outer_launches", C.c_uint64), ("outer_seconds", C.c_double), ("outer_bytes", C.c_uint64),
                ("dofs", C.c_uint64)]


EXPORTS = (
    "fem_options_default", "fem_hex_element", "fem_hex_stress", "fem_solver_new_box", "fem_solver_solve", "fem_solver_solve_device",
    "fem_solver_solve_many", "fem_solver_solve_many_device",
    "fem_solver_method", "fem_solver_set_max_iter", "fem_solver_set_tol", "fem_solver_free",
    "fem_solver_free_count", "fem_solver_unknown_count", "fem_solver_free_list", "fem_solver_point_grid",
    "fem_solver_rhs", "fem_solver_apply", "fem_solver_precondition", "fem_solver_level_info",
    "fem_solver_device_name", "fem_model_new_box", "fem_model_point_count", "fem_model_unknown_count",
    "fem_model_free_count", "fem_model_free_list", "fem_model_point_grid", "fem_model_load", "fem_model_free",
    "fem_model_tet_cells",
    "fem_solver_set_profile", "fem_solver_profile", "fem_solver_time_kernels",
    "fem_server_new", "fem_server_submit", "fem_server_wait", "fem_server_cancel", "fem_server_free",
    "fem_solver_new_csr", "fem_solver_matrix", "fem_dense_solve_device",
    "fem_dense_solve_device_noalloc", "fem_solver_crossover_counts", "fem_server_memory", "fem_server_batching",
    "fem_assemble", "fem_assembly_stats_of", "fem_assembly_download", "fem_assembly_device", "fem_assembly_free",
)

_lib = None


def build(force: bool = False) -> None:
    """Compile libfemahoy_amd.so for gfx950 with hipcc (make)."""
    cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else []) + ["libfemahoy_amd.so"]
    subprocess.run(cmd, check=True)


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FemError(f"{LIB_PATH} is not built: run `make -C {HERE}` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P, SZ, ERR = C.c_void_p, C.c_size_t, C.c_char_p
    F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    sig = {
        "fem_options_default": (None, [C.POINTER(_Options)]),
        "fem_solver_new_box": (C.c_int, [C.POINTER(_Box), C.POINTER(_Options), C.POINTER(P), ERR, SZ]),
        "fem_solver_solve": (C.c_int, [P, F64P, F64P, CANCEL_FN, P, C.POINTER(_Stats), ERR, SZ]),
        "fem_solver_solve_device": (C.c_int, [P, P, P, CANCEL_FN, P, C.POINTER(_Stats), ERR, SZ]),
        "fem_solver_solve_many": (C.c_int, [P, C.c_uint64, F64P, F64P, CANCEL_FN, P, P, ERR, SZ]),
        "fem_solver_solve_many_device": (C.c_int, [P, C.c_uint64, P, P, CANCEL_FN, P, P, ERR, SZ]),
        "fem_solver_method": (C.c_int, [P]),
        "fem_solver_set_max_iter": (None, [P, C.c_uint64]),
        "fem_solver_set_tol": (None, [P, C.c_double]),
        "fem_solver_free": (None, [P]),
        "fem_solver_free_count": (C.c_uint64, [P]),
        "fem_solver_unknown_count": (C.c_uint64, [P]),
        "fem_solver_free_list": (C.c_int, [P, I64P]),
        "fem_solver_point_grid": (C.c_int, [P, U32P]),
        "fem_solver_rhs": (C.c_int, [P, F64P, ERR, SZ]),
        "fem_solver_apply": (C.c_int, [P, F64P, F64P, ERR, SZ]),
        "fem_solver_precondition": (C.c_int, [P, F64P, F64P, ERR, SZ]),
        "fem_solver_level_info": (C.c_int, [P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
        "fem_solver_device_name": (C.c_char_p, [P]),
        "fem_model_new_box": (C.c_int, [C.POINTER(_Box), C.POINTER(P), ERR, SZ]),
        "fem_model_point_count": (C.c_uint64, [P]),
        "fem_model_unknown_count": (C.c_uint64, [P]),
        "fem_model_free_count": (C.c_uint64, [P]),
        "fem_model_free_list": (C.c_int, [P, I64P]),
        "fem_model_point_grid": (C.c_int, [P, U32P]),
        "fem_model_load": (C.c_int, [P, F64P]),
        "fem_model_tet_cells": (C.c_int, [P, U32P, np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]),
        "fem_model_free": (None, [P]),
        "fem_solver_set_profile": (None, [P, C.c_int]),
        "fem_solver_profile": (C.c_int, [P, C.POINTER(Profile)]),
        "fem_solver_time_kernels": (C.c_int, [P, C.c_int, C.POINTER(Profile)]),
        "fem_server_new": (C.c_int, [C.POINTER(_ServerOptions), C.POINTER(P), ERR, SZ]),
        "fem_server_submit": (C.c_int, [P, C.POINTER(_Box), C.POINTER(_Options), P, C.POINTER(C.c_uint64), ERR, SZ]),
        "fem_server_wait": (C.c_int, [P, C.c_uint64, C.c_double, P, C.c_uint64, C.POINTER(_Stats),
                                      C.POINTER(RequestInfo), ERR, SZ]),
        "fem_server_cancel": (C.c_int, [P, C.c_uint64]),
        "fem_server_free": (None, [P]),
        "fem_server_memory": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "fem_server_batching": (C.c_int, [P, C.POINTER(BatchStats)]),
        "fem_solver_matrix": (C.c_int, [P, C.POINTER(C.c_uint64), P, P, P, ERR, SZ]),
        "fem_dense_solve_device": (C.c_int, [C.c_int, P, P, P, ERR, SZ]),
        "fem_dense_solve_device_noalloc": (C.c_int, [C.c_int, P, P, P, ERR, SZ]),
        "fem_solver_crossover_counts": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
        "fem_test_symbolic": (C.c_int, [C.POINTER(_Csr), P, P, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), ERR, SZ]),
        "fem_solver_new_csr": (C.c_int, [C.POINTER(_Csr), C.POINTER(_Unknowns), C.POINTER(_Options), C.POINTER(P),
                                         ERR, SZ]),
        "fem_test_csr_levels_new": (C.c_int, [C.POINTER(_Csr), C.POINTER(_GridLevel), C.c_void_p, C.c_uint64, C.c_int,
                                              C.POINTER(P), ERR, SZ]),
        "fem_test_csr_levels_count": (C.c_int, [P]),
        "fem_test_csr_level_sizes": (C.c_int, [P, C.c_int] + [C.POINTER(C.c_uint64)] * 4),
        "fem_test_csr_level_arrays": (C.c_int, [P, C.c_int, P, P, P, P, P, P, P, P]),
        "fem_test_csr_levels_free": (None, [P]),
        "fem_assemble": (C.c_int, [C.POINTER(_Mesh), C.c_int, C.POINTER(P), ERR, SZ]),
        "fem_assembly_stats_of": (None, [P, C.POINTER(AssemblyStats)]),
        "fem_assembly_download": (C.c_int, [P, P, P, P, P, P, ERR, SZ]),
        "fem_assembly_device": (None, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "fem_assembly_free": (None, [P]),
        "fem_test_assembly_pattern": (C.c_int, [C.POINTER(_Mesh), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), P, P,
                                                P, ERR, SZ]),
        "fem_test_hex_element": (C.c_int, [F64P, C.c_double, C.c_double, F64P, ERR, SZ]),
        "fem_hex_element": (C.c_int, [F64P, F64P, F64P, F64P, F64P, ERR, SZ]),
        "fem_hex_stress": (C.c_int, [F64P, F64P, F64P, F64P, F64P, F64P, F64P, F64P, ERR, SZ]),
        "fem_test_supernodal": (C.c_int, [C.POINTER(_Csr), P, C.c_uint64, C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint64), P, P, P, P, P, ERR, SZ]),
        "fem_test_hex_box_system": (C.c_int, [C.POINTER(_Box), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), P, P, P,
                                              P, C.POINTER(C.c_double), ERR, SZ]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _err():
    return C.create_string_buffer(2048)


@dataclass
class Box:
    """A structured box model: `n` cells on [0, length], one material value per cell (hex8) or per
    tetrahedron (6 per cell, in the reference's numbering of tetrahedra)."""

    kind: str
    n: Sequence[int]
    length: Sequence[float]
    coef0: np.ndarray
    coef1: Optional[np.ndarray] = None
    fixed: dict = field(default_factory=dict)       # heat: face -> temperature; elasticity: face -> True (clamped)
    source: float = 0.0
    body: Sequence[float] = (0.0, 0.0, 0.0)
    traction: dict = field(default_factory=dict)    # face -> (tx, ty, tz)

    def cells(self) -> int:
        return int(np.prod(self.n))

    def _struct(self):
        b = _Box()
        b.kind = KINDS[self.kind]
        for d in range(3):
            b.n[d] = int(self.n[d])
            b.length[d] = float(self.length[d])
        count = self.cells() * (1 if self.kind in HEX_KINDS else 6)
        c0 = np.ascontiguousarray(np.broadcast_to(np.asarray(self.coef0, np.float64), (count,)))
        keep = [c0]
        b.coef0 = c0.ctypes.data_as(C.POINTER(C.c_double))
        if self.coef1 is not None:
            c1 = np.ascontiguousarray(np.broadcast_to(np.asarray(self.coef1, np.float64), (count,)))
            keep.append(c1)
            b.coef1 = c1.ctypes.data_as(C.POINTER(C.c_double))
        for f, v in self.fixed.items():
            i = FACES.index(f)
            b.fixed[i] = 1
            b.value[i] = 0.0 if v is True else float(v)
        b.source = float(self.source)
        for d in range(3):
            b.body[d] = float(self.body[d])
        for f, t in self.traction.items():
            i = FACES.index(f)
            for d in range(3):
                b.traction[3 * i + d] = float(t[d])
        return b, keep


@dataclass
class Stats:
    """solve.rs:81-86 Stats (iterations, residual) and what the device solve adds."""

    iterations: int
    residual: float
    true_residual: float
    seconds: float
    setup_seconds: float
    levels: int
    coarse_unknowns: int
    restarts: int
    method: str
    precision: str
    hbm_bytes: int = 0  # algorithmic bytes of the solve's launches (box solvers)


class Model:
    """The numbering, free unknowns and loads of a box model (host only, no device)."""

    def __init__(self, box: Box):
        b, keep = box._struct()
        h = C.c_void_p()
        err = _err()
        if lib().fem_model_new_box(C.byref(b), C.byref(h), err, 2048):
            raise FemError(err.value.decode())
        self.h = h
        self._n, self._length = list(box.n), list(box.length)
        del keep

    def __del__(self):
        if getattr(self, "h", None):
            lib().fem_model_free(self.h)
            self.h = None

    def points(self) -> int:
        return lib().fem_model_point_count(self.h)

    def free_list(self) -> np.ndarray:
        out = np.zeros(lib().fem_model_free_count(self.h), np.int64)
        lib().fem_model_free_list(self.h, out)
        return out

    def point_grid(self) -> np.ndarray:
        out = np.zeros(3 * self.points(), np.uint32)
        lib().fem_model_point_grid(self.h, out)
        return out.reshape(-1, 3)

    def load(self) -> np.ndarray:
        out = np.zeros(lib().fem_model_unknown_count(self.h))
        lib().fem_model_load(self.h, out)
        return out

    def tet_cells(self):
        """The cell (x fastest) and Kuhn type of each tetrahedron, in the reference's numbering."""
        ntet = 6 * int(np.prod(self._n))
        cell = np.zeros(ntet, np.uint32)
        kind = np.zeros(ntet, np.uint8)
        if lib().fem_model_tet_cells(self.h, cell, kind):
            raise FemError("a hex8 model has no tetrahedra")
        return cell, kind

    def tet_centroids(self) -> np.ndarray:
        """The centroid of each tetrahedron (reference numbering), shape (tets, 3). A Kuhn chain
        of axes (a, b, c) from the cell's lowest corner has its centroid at (3/4, 2/4, 1/4) of the
        cell along a, b, c."""
        cell, kind = self.tet_cells()
        n = np.asarray(self._n, np.int64)
        h = np.asarray(self._length, np.float64) / n
        ijk = np.stack([cell % n[0], (cell // n[0]) % n[1], cell // (n[0] * n[1])], axis=1).astype(np.float64)
        frac = np.zeros((6, 3))
        for t, axes in enumerate(KUHN_AXES):
            for r, a in enumerate(axes):
                frac[t, a] = (3 - r) / 4.0
        return (ijk + frac[kind]) * h


def box_aspect(box: "Box") -> float:
    """The largest ratio of a cell's longest side to its shortest (elastic.rs:525-536), the cell
    sides being the differences of the planes h i (h = length / n) as the library makes them."""
    sides = []
    for d in range(3):
        h = float(box.length[d]) / float(box.n[d])
        pl = np.arange(int(box.n[d]) + 1, dtype=np.float64) * h
        sides.append(np.unique(np.diff(pl)))
    g = np.stack(np.meshgrid(*sides, indexing="ij"), axis=-1).reshape(-1, 3)
    return float(np.max(g.max(axis=1) / g.min(axis=1)))


@dataclass
class Flops:
    """Crossover::Flops (solve.rs:28-39): direct when the nested-dissection factor has at most
    `entries` entries and (`flat`, or at most `flops` flops per unknown). A plain int crossover is
    Crossover::Unknowns."""
    flops: float = 1.5e5          # elastic.rs:263 DIRECT_FLOPS
    entries: int = 250_000_000    # elastic.rs:269 DIRECT_ENTRIES
    flat: bool = False


def _options(method=None, precision="fp32", tol=PCG_TOL, max_iter=PCG_MAX_ITER, crossover=STEADY_DIRECT_LIMIT,
             levels=0, cheb_degree=0, cheb_lower=0.1, cheb_upper=1.1, power_steps=20, coarse_max=6000, refine=True,
             device=0, coarse_galerkin=True, direct_solver="auto") -> "_Options":
    o = _Options()
    lib().fem_options_default(C.byref(o))
    o.tol, o.max_iter, o.method, o.precision = tol, max_iter, METHODS[method], PRECISIONS[precision]
    if isinstance(crossover, Flops):
        o.crossover_kind, o.crossover_flat = 1, int(crossover.flat)
        o.crossover_flops, o.crossover_entries = float(crossover.flops), int(crossover.entries)
    else:
        o.crossover_kind, o.crossover = 0, int(crossover)
    o.levels, o.cheb_degree = levels, cheb_degree
    o.cheb_lower, o.cheb_upper, o.power_steps = cheb_lower, cheb_upper, power_steps
    o.coarse_max, o.refine, o.device = coarse_max, int(refine), device
    o.coarse_galerkin = int(coarse_galerkin)
    if direct_solver not in DIRECT_SOLVERS:
        raise FemError(f"unknown direct solver {direct_solver}: auto, dense or multifrontal")
    o.direct_solver = DIRECT_SOLVERS[direct_solver]
    return o


class Solver:
    """Solver::new / solve / method / with_max_iter over the device library."""

    def __init__(self, box: Box, method: Optional[str] = None, precision: str = "fp32", tol: float = PCG_TOL,
                 max_iter: int = PCG_MAX_ITER, crossover=None, levels: int = 0,
                 cheb_degree: int = 0, cheb_lower: float = 0.1, cheb_upper: float = 1.1, power_steps: int = 20,
                 coarse_max: int = 6000, refine: bool = True, device: int = 0, coarse_galerkin: bool = True,
                 direct_solver: str = "auto"):
        if crossover is None:
            # the box-grid elasticity caller's choice (elastic.rs:525-547): Crossover::Flops, flat when
            # a cell's longest side is more than MULTIGRID_ASPECT times its shortest; else Unknowns
            crossover = (Flops(flat=box_aspect(box) > MULTIGRID_ASPECT) if box.kind == "elastic_hex8"
                         else STEADY_DIRECT_LIMIT)
        o = _options(method, precision, tol, max_iter, crossover, levels, cheb_degree, cheb_lower, cheb_upper,
                     power_steps, coarse_max, refine, device, coarse_galerkin, direct_solver)
        b, keep = box._struct()
        h = C.c_void_p()
        err = _err()
        if lib().fem_solver_new_box(C.byref(b), C.byref(o), C.byref(h), err, 2048):
            raise FemError(err.value.decode())
        del keep
        self.h = h
        self.n = int(lib().fem_solver_free_count(self.h))

    new = classmethod(lambda cls, *a, **k: cls(*a, **k))

    def __del__(self):
        if getattr(self, "h", None):
            lib().fem_solver_free(self.h)
            self.h = None

    def crossover_counts(self):
        """(entries, flops) of the nested-dissection factor that a FLOPS crossover compared, or
        None when the solver's choice did not analyse the matrix."""
        e, f = C.c_uint64(), C.c_double()
        if lib().fem_solver_crossover_counts(self.h, C.byref(e), C.byref(f)):
            return None
        return int(e.value), float(f.value)

    def method(self) -> str:
        return ("direct", "multigrid")[lib().fem_solver_method(self.h)]

    def with_max_iter(self, max_iter: int) -> "Solver":
        lib().fem_solver_set_max_iter(self.h, max_iter)
        return self

    def set_tol(self, tol: float) -> "Solver":
        lib().fem_solver_set_tol(self.h, tol)
        return self

    def free_list(self) -> np.ndarray:
        out = np.zeros(self.n, np.int64)
        lib().fem_solver_free_list(self.h, out)
        return out

    def rhs(self) -> np.ndarray:
        out = np.zeros(self.n)
        err = _err()
        if lib().fem_solver_rhs(self.h, out, err, 2048):
            raise FemError(err.value.decode())
        return out

    def apply(self, x) -> np.ndarray:
        y = np.zeros(self.n)
        err = _err()
        if lib().fem_solver_apply(self.h, np.ascontiguousarray(x, np.float64), y, err, 2048):
            raise FemError(err.value.decode())
        return y

    def matrix(self):
        """The reduced matrix of the model (fem_solver_matrix): (ptr, idx, val) of a CSR over the
        free unknowns in reference order, as the reference's callers assemble it."""
        nnz = C.c_uint64()
        err = _err()
        L = lib()
        if L.fem_solver_matrix(self.h, C.byref(nnz), None, None, None, err, 2048):
            raise FemError(err.value.decode())
        ptr = np.zeros(self.n + 1, np.uint64)
        idx = np.zeros(nnz.value, np.uint32)
        val = np.zeros(nnz.value)
        vp = lambda a: C.c_void_p(a.ctypes.data)
        if L.fem_solver_matrix(self.h, C.byref(nnz), vp(ptr), vp(idx), vp(val), err, 2048):
            raise FemError(err.value.decode())
        return ptr, idx, val

    def precondition(self, r) -> np.ndarray:
        z = np.zeros(self.n)
        err = _err()
        if lib().fem_solver_precondition(self.h, np.ascontiguousarray(r, np.float64), z, err, 2048):
            raise FemError(err.value.decode())
        return z

    def level_info(self, level: int):
        n = C.c_uint64()
        lm = C.c_double()
        if lib().fem_solver_level_info(self.h, level, C.byref(n), C.byref(lm)):
            raise FemError("no such level")
        return int(n.value), float(lm.value)

    def device_name(self) -> str:
        return lib().fem_solver_device_name(self.h).decode()

    def set_profile(self, on: bool) -> None:
        """HIP events around the finest-level smoother and outer operator launches of later
        solves (eager launches while on)."""
        lib().fem_solver_set_profile(self.h, 1 if on else 0)

    def profile(self) -> Profile:
        p = Profile()
        if lib().fem_solver_profile(self.h, C.byref(p)):
            raise FemError("no profile")
        return p

    def time_kernels(self, reps: int = 100) -> Profile:
        """Mean device time of the finest smoother step and of the outer q = A p, each repeated
        `reps` times back to back between one pair of HIP events on the solver's stream."""
        p = Profile()
        if lib().fem_solver_time_kernels(self.h, int(reps), C.byref(p)):
            raise FemError("kernel timing failed")
        return p

    def solve(self, b, x=None, cancelled: Optional[Callable[[], bool]] = None):
        """Solve A x = b; `x` is the start value (zeros if None). Returns (x, Stats)."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x is None else np.array(x, np.float64, copy=True)
        if b.shape != (self.n,) or x.shape != (self.n,):
            raise FemError(f"b and x need {self.n} values")
        cb = CANCEL_FN(lambda _ctx: 1 if cancelled() else 0) if cancelled else CANCEL_FN()
        st = _Stats()
        err = _err()
        rc = lib().fem_solver_solve(self.h, b, x, cb, None, C.byref(st), err, 2048)
        if rc:
            raise FemError(err.value.decode())
        return x, _stats(st)

    def solve_many(self, B, X=None, cancelled: Optional[Callable[[], bool]] = None):
        """Solve A x_i = b_i for the k rows of B (k x n) against this one setup
        (fem_solver_solve_many: the k solves run together, their iterations batched); X holds the
        start values (zeros if None). Returns (X, [Stats]); each row equals solve() of that right
        side. Raises the first right side's error after all have ended."""
        B = np.ascontiguousarray(B, np.float64)
        if B.ndim != 2 or B.shape[1] != self.n:
            raise FemError(f"B needs k rows of {self.n} values")
        k = B.shape[0]
        X = np.zeros_like(B) if X is None else np.array(X, np.float64, copy=True, order="C")
        if X.shape != B.shape:
            raise FemError("X needs the shape of B")
        cb = CANCEL_FN(lambda _ctx: 1 if cancelled() else 0) if cancelled else CANCEL_FN()
        st = (_Stats * max(k, 1))()
        err = _err()
        rc = lib().fem_solver_solve_many(self.h, k, B.reshape(-1), X.reshape(-1), cb, None, C.cast(st, C.c_void_p), err,
                                         2048)
        if rc:
            raise FemError(err.value.decode())
        return X, [_stats(st[i]) for i in range(k)]

    def solve_device(self, d_b: int, d_x: int, cancelled: Optional[Callable[[], bool]] = None) -> Stats:
        """Solve with b and x already in device memory (raw HIP pointers)."""
        cb = CANCEL_FN(lambda _ctx: 1 if cancelled() else 0) if cancelled else CANCEL_FN()
        st = _Stats()
        err = _err()
        if lib().fem_solver_solve_device(self.h, C.c_void_p(d_b), C.c_void_p(d_x), cb, None, C.byref(st), err, 2048):
            raise FemError(err.value.decode())
        return _stats(st)


@dataclass
class Grid:
    """The unknowns of a grid level (multigrid.rs:55-59 Level): the plane coordinates of each axis,
    and the grid point (plane indices) and component of each unknown."""

    planes: Sequence[np.ndarray]
    points: np.ndarray  # (n, 3) plane indices
    comp: np.ndarray    # (n,) 0 for a scalar, 0..2 for a displacement

    def _struct(self, n: int):
        g = _GridLevel()
        keep = []
        for d in range(3):
            pl = np.ascontiguousarray(self.planes[d], np.float64)
            keep.append(pl)
            g.planes[d] = pl.ctypes.data_as(C.POINTER(C.c_double))
            g.nplanes[d] = pl.size
        pts = np.ascontiguousarray(self.points, np.uint32).reshape(-1)
        cmp = np.ascontiguousarray(self.comp, np.uint8).reshape(-1)
        if pts.size != 3 * n or cmp.size != n:
            raise FemError(f"the grid level needs a point and a component for each of {n} unknowns")
        keep += [pts, cmp]
        g.points = pts.ctypes.data_as(C.POINTER(C.c_uint32))
        g.comp = cmp.ctypes.data_as(C.POINTER(C.c_uint8))
        return g, keep


def _csr_struct(ptr, idx, val, rows: int, cols: int):
    ptr = np.ascontiguousarray(ptr, np.uint64)
    idx = np.ascontiguousarray(idx, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    if ptr.size != rows + 1 or idx.size != int(ptr[-1]) or val.size != int(ptr[-1]):
        raise FemError("a CSR matrix needs rows + 1 offsets and ptr[rows] columns and values")
    m = _Csr()
    m.rows, m.cols = rows, cols
    m.ptr = ptr.ctypes.data_as(C.POINTER(C.c_uint64))
    m.idx = idx.ctypes.data_as(C.POINTER(C.c_uint32))
    m.val = val.ctypes.data_as(C.POINTER(C.c_double))
    return m, [ptr, idx, val]


class CsrSolver(Solver):
    """Solver::new(a, unknowns, method, crossover, None) (solve.rs:92-183) for an assembled matrix
    of the free unknowns (CSR, both triangles, sorted rows): `unknowns` is "grid" (with `grid`),
    "points", or "quadratic" (with the prolongation `p` = (ptr, idx, val, cols) and, optionally,
    `coarse_grid`, the grid level of the linear unknowns). b and x are in the matrix's order."""

    def __init__(self, ptr, idx, val, unknowns: str = "grid", grid: Optional[Grid] = None, p=None,
                 coarse_grid: Optional[Grid] = None, method: Optional[str] = None, precision: str = "fp32",
                 tol: float = PCG_TOL, max_iter: int = PCG_MAX_ITER, crossover=STEADY_DIRECT_LIMIT,
                 levels: int = 0, cheb_degree: int = 0, cheb_lower: float = 0.1, cheb_upper: float = 1.1,
                 power_steps: int = 20, coarse_max: int = 6000, refine: bool = True, device: int = 0,
                 xyz=None, cells=None, direct_solver: str = "auto"):
        n = len(ptr) - 1
        o = _options(method, precision, tol, max_iter, crossover, levels, cheb_degree, cheb_lower, cheb_upper,
                     power_steps, coarse_max, refine, device, True)
        o.direct_solver = DIRECT_SOLVERS[direct_solver]
        a, keep = _csr_struct(ptr, idx, val, n, n)
        u = _Unknowns()
        u.kind = UNKNOWNS[unknowns]
        if unknowns == "grid":
            if grid is None:
                raise FemError("grid unknowns need their grid level")
            u.grid, k = grid._struct(n)
            keep += k
        elif unknowns == "quadratic":
            if p is None:
                raise FemError("quadratic unknowns need the prolongation")
            pptr, pidx, pval, pcols = p
            u.p, k = _csr_struct(pptr, pidx, pval, n, int(pcols))
            keep += k
            if coarse_grid is not None:
                u.coarse_has_grid = 1
                u.coarse_grid, k = coarse_grid._struct(int(pcols))
                keep += k
        if xyz is not None:
            xyz = np.ascontiguousarray(xyz, np.float64).reshape(n, 3)
            u.xyz = xyz.ctypes.data
            keep.append(xyz)
        if cells is not None:
            origin, element = cells
            origin = np.ascontiguousarray(origin, np.uint32).reshape(-1, 3)
            element = np.ascontiguousarray(element, np.float64).reshape(-1, 576)
            cs = _Cells(len(origin), origin.ctypes.data, element.ctypes.data)
            u.cells = C.pointer(cs)
            keep += [origin, element, cs]
        h = C.c_void_p()
        err = _err()
        if lib().fem_solver_new_csr(C.byref(a), C.byref(u), C.byref(o), C.byref(h), err, 2048):
            raise FemError(err.value.decode())
        del keep
        self.h = h
        self.n = int(lib().fem_solver_free_count(self.h))


def symbolic(ptr, idx, xyz):
    """The library's host nested dissection (order.rs:19-53) and symbolic counts (symbolic.rs) of a
    square pattern, as the FLOPS crossover computes them: (order, entries, flops, largest_front,
    supernodes). No device is touched."""
    n = len(ptr) - 1
    a, keep = _csr_struct(ptr, idx, np.zeros(int(ptr[-1])), n, n)
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(n, 3)
    order = np.zeros(n, np.uint64)
    e, f, lf, sn = C.c_uint64(), C.c_double(), C.c_uint64(), C.c_uint64()
    err = _err()
    if lib().fem_test_symbolic(C.byref(a), xyz.ctypes.data, order.ctypes.data, C.byref(e), C.byref(f), C.byref(lf),
                               C.byref(sn), err, 2048):
        raise FemError(err.value.decode())
    del keep
    return order.astype(np.int64), int(e.value), float(f.value), int(lf.value), int(sn.value)


def supernodal(ptr, idx, xyz, join: int = 0):
    """The library's host structure of the multifrontal factor of a square pattern (symbolic.rs:
    41-150 in the nested-dissection order of xyz): (perm, start, rptr, rows, sparent). No device."""
    n = len(ptr) - 1
    a, keep = _csr_struct(ptr, idx, np.zeros(int(ptr[-1])), n, n)
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(n, 3)
    ns, nr = C.c_uint64(), C.c_uint64()
    err = _err()
    L = lib()
    if L.fem_test_supernodal(C.byref(a), xyz.ctypes.data, int(join), C.byref(ns), C.byref(nr), None, None, None, None,
                             None,
                             err, 2048):
        raise FemError(err.value.decode())
    perm = np.zeros(n, np.uint64)
    start, rptr = np.zeros(ns.value + 1, np.uint64), np.zeros(ns.value + 1, np.uint64)
    rows, sparent = np.zeros(max(nr.value, 1), np.uint64), np.zeros(max(ns.value, 1), np.int64)
    if L.fem_test_supernodal(C.byref(a), xyz.ctypes.data, int(join), C.byref(ns), C.byref(nr), perm.ctypes.data,
                             start.ctypes.data, rptr.ctypes.data, rows.ctypes.data, sparent.ctypes.data, err, 2048):
        raise FemError(err.value.decode())
    del keep
    return (perm.astype(np.int64), start.astype(np.int64), rptr.astype(np.int64), rows[:nr.value].astype(np.int64),
            sparent[:ns.value])


def csr_levels(ptr, idx, val, grid: Grid, coarsest: int = 3000, max_levels: int = 0, cells=None):
    """The host setup of the assembled grid multigrid alone (no device): the reference's level loop
    with coarsen and Galerkin products. Returns a list of dicts per level: n, a = (ptr, idx, val),
    p = (ptr, idx, val, cols) from the next level (None on the last), points, comp."""
    n = len(ptr) - 1
    a, keep = _csr_struct(ptr, idx, val, n, n)
    g, k = grid._struct(n)
    keep += k
    h = C.c_void_p()
    err = _err()
    L = lib()
    cp = None
    if cells is not None:
        origin, element = cells
        origin = np.ascontiguousarray(origin, np.uint32).reshape(-1, 3)
        element = np.ascontiguousarray(element, np.float64).reshape(-1, 576)
        cs = _Cells(len(origin), origin.ctypes.data, element.ctypes.data)
        keep += [origin, element, cs]
        cp = C.cast(C.pointer(cs), C.c_void_p)
    if L.fem_test_csr_levels_new(C.byref(a), C.byref(g), cp, coarsest, max_levels, C.byref(h), err, 2048):
        raise FemError(err.value.decode())
    try:
        out = []
        for l in range(L.fem_test_csr_levels_count(h)):
            sz = [C.c_uint64() for _ in range(4)]
            L.fem_test_csr_level_sizes(h, l, *[C.byref(v) for v in sz])
            nl, annz, pcols, pnnz = [int(v.value) for v in sz]
            ap, ai, av = np.zeros(nl + 1, np.uint64), np.zeros(annz, np.uint32), np.zeros(annz)
            last = pcols == 0 and pnnz == 0 and l == L.fem_test_csr_levels_count(h) - 1
            pp, pi, pv = np.zeros(nl + 1, np.uint64), np.zeros(pnnz, np.uint32), np.zeros(pnnz)
            pts, cmp = np.zeros(3 * nl, np.uint32), np.zeros(nl, np.uint8)
            vp = lambda x: C.c_void_p(x.ctypes.data)
            L.fem_test_csr_level_arrays(h, l, vp(ap), vp(ai), vp(av), vp(pp), vp(pi), vp(pv), vp(pts), vp(cmp))
            out.append(dict(n=nl, a=(ap, ai, av), p=None if last else (pp, pi, pv, pcols), points=pts.reshape(-1, 3),
                            comp=cmp))
        return out
    finally:
        L.fem_test_csr_levels_free(h)


def _stats(st: _Stats) -> Stats:
    return Stats(int(st.iterations), float(st.residual), float(st.true_residual), float(st.seconds),
                 float(st.setup_seconds), int(st.levels), int(st.coarse_unknowns), int(st.restarts),
                  ("direct", "multigrid")[st.method] if st.method in (0, 1) else "?",
                 {v: k for k, v in PRECISIONS.items()}.get(st.precision, "?"), int(st.hbm_bytes))


class DeviceArray:
    """A float64 array in device memory (hipMalloc through the HIP runtime), for the entry points
    that take device pointers."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so") if not os.path.exists("/opt/rocm/lib/libamdhip64.so") \
                else C.CDLL("/opt/rocm/lib/libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.shape, self.nbytes = host.shape, host.nbytes
        p = C.c_void_p()
        if self.hip().hipMalloc(C.byref(p), max(self.nbytes, 8)):
            raise FemError("hipMalloc failed")
        self.ptr = p.value
        if self.nbytes and self.hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1):
            raise FemError("hipMemcpy to the device failed")

    def get(self) -> np.ndarray:
        out = np.empty(self.shape)
        if self.nbytes and self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2):
            raise FemError("hipMemcpy from the device failed")
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def dense_solve(a, b, entry: str = "fem_dense_solve_device") -> np.ndarray:
    """x = A^-1 b for a dense SPD matrix through the device's dense LDL^T (fem_dense_solve_device:
    Cholesky::factor and Cholesky::solve, crates/fem/src/cholesky.rs:44-160, for a dense matrix).
    The arrays are copied to device memory for the call; `entry` names the C entry point."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = b.shape[0]
    if a.shape != (n, n):
        raise FemError("the matrix must be n x n for a right side of n values")
    da, db, dx = DeviceArray(a), DeviceArray(b), DeviceArray(np.zeros(n))
    err = _err()
    if getattr(lib(), entry)(n, da.ptr, db.ptr, dx.ptr, err, len(err)) != 0:
        raise FemError(err.value.decode())
    return dx.get()


MESH_KINDS = {"heat_hex8": 0, "heat_tet4": 1, "elastic_tet10": 2, "elastic_hex8": 3}


@dataclass
class Mesh:
    """A mesh for fem_assemble: points (n x 3), cells (c x 8, 4 or 10 point numbers), a material
    value per cell (conductivity, or Young's modulus with `coef1` Poisson's ratio), the held
    unknowns and their values (one per unknown: a point, or 3 point + component), a heat source or
    body force per volume, and optionally a further load per unknown added after them."""

    kind: str
    xyz: np.ndarray
    cells: np.ndarray
    coef0: np.ndarray
    fixed: np.ndarray
    value: np.ndarray
    coef1: Optional[np.ndarray] = None
    source: float = 0.0
    body: Sequence[float] = (0.0, 0.0, 0.0)
    load: Optional[np.ndarray] = None
    ortho: Optional[np.ndarray] = None  # elasticity: per cell the 9 constants of material9, for coef0 / coef1

    def _struct(self):
        if self.kind not in MESH_KINDS:
            raise FemError(f"unknown mesh kind {self.kind!r}")
        keep = {
            "xyz": np.ascontiguousarray(self.xyz, np.float64).reshape(-1),
            "conn": np.ascontiguousarray(self.cells, np.uint32).reshape(-1),
            "fixed": np.ascontiguousarray(self.fixed, np.uint8).reshape(-1),
            "value": np.ascontiguousarray(self.value, np.float64).reshape(-1),
        }
        if self.coef0 is not None:
            keep["coef0"] = np.ascontiguousarray(self.coef0, np.float64).reshape(-1)
        if self.ortho is not None:
            keep["ortho"] = np.ascontiguousarray(self.ortho, np.float64).reshape(-1)
            if keep["ortho"].size != 9 * np.asarray(self.cells).shape[0]:
                raise FemError("ortho needs 9 constants per cell")
        if self.coef1 is not None:
            keep["coef1"] = np.ascontiguousarray(self.coef1, np.float64).reshape(-1)
        if self.load is not None:
            keep["load"] = np.ascontiguousarray(self.load, np.float64).reshape(-1)
        m = _Mesh()
        m.kind = MESH_KINDS[self.kind]
        m.points = keep["xyz"].size // 3
        m.cells = np.asarray(self.cells).shape[0]
        for k, v in keep.items():
            setattr(m, k, v.ctypes.data)
        m.source = float(self.source)
        for d in range(3):
            m.body[d] = float(self.body[d])
        units = m.points * (3 if self.kind.startswith("elastic") else 1)
        if keep["fixed"].size != units or keep["value"].size != units:
            raise FemError(f"fixed and value need one entry per unknown ({units})")
        if "load" in keep and keep["load"].size != units:
            raise FemError(f"load needs one entry per unknown ({units})")
        return m, keep


class Assembly:
    """The system of a mesh assembled on the device (fem_assemble): the reduced CSR matrix over the
    free unknowns (increasing) and its right side."""

    def __init__(self, mesh: Mesh, device: int = 0):
        m, self._keep = mesh._struct()
        h = C.c_void_p()
        err = _err()
        if lib().fem_assemble(C.byref(m), int(device), C.byref(h), err, 2048):
            raise FemError(err.value.decode())
        self.h = h
        st = AssemblyStats()
        lib().fem_assembly_stats_of(self.h, C.byref(st))
        self.stats = st
        self.n = int(st.free)

    def arrays(self):
        """(ptr, idx, val, rhs, free) of the reduced system, on the host."""
        n, ne = self.n, int(self.stats.entries)
        ptr = np.zeros(n + 1, np.uint64)
        idx = np.zeros(max(ne, 1), np.uint32)
        val = np.zeros(max(ne, 1))
        rhs = np.zeros(max(n, 1))
        free = np.zeros(max(n, 1), np.int64)
        err = _err()
        if lib().fem_assembly_download(self.h, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, rhs.ctypes.data,
                                       free.ctypes.data, err, 2048):
            raise FemError(err.value.decode())
        return ptr, idx[:ne], val[:ne], rhs[:n], free[:n]

    def device(self):
        """Device pointers (ptr, idx, val, rhs) of the reduced system, valid while this lives."""
        out = [C.c_void_p() for _ in range(4)]
        lib().fem_assembly_device(self.h, *[C.byref(o) for o in out])
        return tuple(o.value for o in out)

    def close(self):
        if getattr(self, "h", None):
            lib().fem_assembly_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def assembly_pattern(mesh: Mesh):
    """The host part of fem_assemble alone (no device): (ptr, idx, free) of the reduced pattern."""
    m, keep = mesh._struct()
    nf, ne = C.c_uint64(), C.c_uint64()
    err = _err()
    if lib().fem_test_assembly_pattern(C.byref(m), C.byref(nf), C.byref(ne), None, None, None, err, 2048):
        raise FemError(err.value.decode())
    ptr = np.zeros(nf.value + 1, np.uint64)
    idx = np.zeros(max(ne.value, 1), np.uint32)
    free = np.zeros(max(nf.value, 1), np.int64)
    if lib().fem_test_assembly_pattern(C.byref(m), C.byref(nf), C.byref(ne), ptr.ctypes.data, idx.ctypes.data,
                                       free.ctypes.data, err, 2048):
        raise FemError(err.value.decode())
    return ptr, idx[:ne.value], free[:nf.value]


def material9(e, nu, g=None) -> np.ndarray:
    """The 9 constants of an orthotropic material with its axes on x, y and z (hex.rs:47-64): E_x,
    E_y, E_z, nu_xy, nu_xz, nu_yz, G_xy, G_xz, G_yz. Scalar e and nu: the isotropic material (E, nu
    and E / (2 (1 + nu)) on every axis); 3 values each: orthotropic, with the shear moduli g."""
    if np.ndim(e) == 0 and np.ndim(nu) == 0 and g is None:
        e, nu = float(e), float(nu)
        gi = e / (2.0 * (1.0 + nu))
        return np.array([e, e, e, nu, nu, nu, gi, gi, gi])
    if g is None or np.size(e) != 3 or np.size(nu) != 3 or np.size(g) != 3:
        raise FemError("an orthotropic material needs 3 Young's moduli, 3 Poisson ratios and 3 shear moduli")
    return np.concatenate([np.ravel(e), np.ravel(nu), np.ravel(g)]).astype(np.float64)


def _alpha3(alpha) -> np.ndarray:
    if alpha is None:
        return np.zeros(3)
    return np.full(3, float(alpha)) if np.ndim(alpha) == 0 else np.ascontiguousarray(alpha, np.float64).reshape(3)


def hex_element(size, e, nu, alpha=None, g=None):
    """The condensed incompatible-mode hex of a box cell (hex.rs:199-279, fem_hex_element), 24 x 24,
    of an isotropic (scalar e, nu) or orthotropic (3 values each, and g) material; with expansion
    coefficients `alpha` (a number, or 3 along x, y and z), also the load of a unit temperature
    change at each corner (24 x 8): (k, load). Host code, no device."""
    k = np.zeros(576)
    load = np.zeros(192)
    err = _err()
    if lib().fem_hex_element(np.ascontiguousarray(size, np.float64), material9(e, nu, g), _alpha3(alpha), k, load, err,
                             2048):
        raise FemError(err.value.decode())
    return k.reshape(24, 24) if alpha is None else (k.reshape(24, 24), load.reshape(24, 8))


def hex_strain_stress(size, e, nu, alpha, xi, u, t, g=None):
    """hex.rs:313-340 strain_stress of the condensed hex (fem_hex_stress): the engineering strain and
    the stress (xx, yy, zz, yz, xz, xy) at the local point xi from the corner displacements u (24)
    and temperature changes t (8). Host code, no device."""
    strain, stress = np.zeros(6), np.zeros(6)
    err = _err()
    if lib().fem_hex_stress(np.ascontiguousarray(size, np.float64), material9(e, nu, g), _alpha3(alpha),
                            np.ascontiguousarray(xi, np.float64), np.ascontiguousarray(u, np.float64),
                            np.ascontiguousarray(t, np.float64), strain, stress, err, 2048):
        raise FemError(err.value.decode())
    return strain, stress


def hex_box_system(box: Box):
    """The reduced system fem_solver_new_box assembles for an elastic_hex8 box, on the host (no
    device): ((ptr, idx, val), rhs, aspect)."""
    b, keep = box._struct()
    n, nnz, aspect = C.c_uint64(), C.c_uint64(), C.c_double()
    err = _err()
    L = lib()
    if L.fem_test_hex_box_system(C.byref(b), C.byref(n), C.byref(nnz), None, None, None, None, C.byref(aspect), err,
                                 2048):
        raise FemError(err.value.decode())
    ptr = np.zeros(n.value + 1, np.uint64)
    idx = np.zeros(max(nnz.value, 1), np.uint32)
    val = np.zeros(max(nnz.value, 1))
    rhs = np.zeros(max(n.value, 1))
    if L.fem_test_hex_box_system(C.byref(b), C.byref(n), C.byref(nnz), ptr.ctypes.data, idx.ctypes.data,
                                 val.ctypes.data, rhs.ctypes.data, C.byref(aspect), err, 2048):
        raise FemError(err.value.decode())
    del keep
    return (ptr, idx[:nnz.value], val[:nnz.value]), rhs[:n.value], aspect.value


class BatchStats(C.Structure):
    """fem_server_batch_stats: the serving loop's batched iterations."""

    _fields_ = [("batches", C.c_uint64), ("iterations", C.c_uint64), ("launches", C.c_uint64),
                ("solo_launches", C.c_uint64), ("graphs", C.c_uint64), ("graph_seconds", C.c_double),
                ("launch_seconds", C.c_double), ("finish_seconds", C.c_double)]


class Server:
    """The serving loop (fem_server_*): solves of uneven sizes admitted by device footprint, set up
    on a thread of their own and advanced together, one PCG iteration per solve at a time; with
    batching (the default) the iterations due go out together, one launch per kernel of each level
    over every running solve."""

    def __init__(self, hbm_budget: int, max_active: int = 8, device: int = 0, setup_threads: int = 2,
                 batched: bool = True):
        o = _ServerOptions(int(hbm_budget), int(max_active), int(device), int(setup_threads), 0 if batched else 1)
        h = C.c_void_p()
        err = _err()
        if lib().fem_server_new(C.byref(o), C.byref(h), err, 2048):
            raise FemError(err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fem_server_free(self.h)
            self.h = None

    __del__ = close

    def submit(self, box: Box, b: Optional[np.ndarray] = None, **options) -> int:
        """Queues a solve of `box` (right side `b` over the free unknowns, or the model's own);
        options as for Solver. Returns the request id."""
        o = _options(**options)
        bs, keep = box._struct()
        ptr = None
        if b is not None:
            b = np.ascontiguousarray(b, np.float64)
            ptr = b.ctypes.data_as(C.c_void_p)
        rid = C.c_uint64()
        err = _err()
        if lib().fem_server_submit(self.h, C.byref(bs), C.byref(o), ptr, C.byref(rid), err, 2048):
            raise FemError(err.value.decode())
        del keep
        return int(rid.value)

    def wait(self, rid: int, n: int, timeout: float = -1.0):
        """The result of request `rid` with `n` free unknowns: (x, Stats, RequestInfo). Raises
        FemError with the solve's error, TimeoutError on timeout."""
        x = np.zeros(n)
        st = _Stats()
        info = RequestInfo()
        err = _err()
        rc = lib().fem_server_wait(self.h, rid, float(timeout), x.ctypes.data_as(C.c_void_p), n, C.byref(st),
                                   C.byref(info), err, 2048)
        if rc == 2:
            raise TimeoutError(f"request {rid} is not done")
        if rc:
            raise FemError(err.value.decode())
        return x, _stats(st), info

    def cancel(self, rid: int) -> bool:
        return lib().fem_server_cancel(self.h, rid) == 0

    def batching(self) -> dict:
        """Counts of the batched iterations (fem_server_batching)."""
        b = BatchStats()
        lib().fem_server_batching(self.h, C.byref(b))
        return {f: (float if t is C.c_double else int)(getattr(b, f)) for f, t in BatchStats._fields_}

    def memory(self):
        """(arena bytes, peak bytes in use, allocations made outside the arena)."""
        a, p, m = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib().fem_server_memory(self.h, C.byref(a), C.byref(p), C.byref(m))
        return int(a.value), int(p.value), int(m.value)



def inclusion_box(n: int, count: int, seed: int, e_ratio: float = 1e3, nu_matrix: float = 0.49,
                  nu_inclusion: float = 0.3, radius=(0.02, 0.05)) -> "tuple[Box, np.ndarray]":
    """Nearly incompressible heterogeneous elasticity: tet10 on n x n/2 x n/2 unit cells (6 Kuhn
    tetrahedra per cell), a matrix of E = 1 and nu = nu_matrix holding `count` spheres of E =
    e_ratio and nu = nu_inclusion, centres uniform in the box and radii uniform in `radius`
    (0.02-0.05 by default) of the box length, from numpy's default_rng(seed); a tetrahedron is in a sphere when its centroid
    is. Clamped at x = 0, traction (1, 0, 0) on x = 1. Returns the box and the inclusion mask
    (reference numbering of tetrahedra)."""
    cells = [n, n // 2, n // 2]
    length = [float(c) for c in cells]
    c = Model(Box("elastic_tet10", cells, length, 1.0, 0.3, fixed={"x0": True})).tet_centroids()
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (count, 3)) * np.asarray(length)
    radii = rng.uniform(radius[0], radius[1], count) * max(length)
    inside = np.zeros(len(c), bool)
    for p, r in zip(centres, radii):
        inside |= ((c - p) ** 2).sum(axis=1) <= r * r
    e = np.where(inside, e_ratio, 1.0)
    nu = np.where(inside, nu_inclusion, nu_matrix)
    box = Box("elastic_tet10", cells, length, e, nu, fixed={"x0": True}, traction={"x1": (1.0, 0.0, 0.0)})
    return box, inside
