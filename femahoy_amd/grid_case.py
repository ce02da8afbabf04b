"""The reference's case files (`grid.toml`) for steady and transient heat conduction and linear
static stress in box parts, read and solved with this library on the device.

The reader follows crates/core/src/grid.rs (`read`, `read_grid`, `read_parts`, `read_faces`, the
loads and probes of `read_heat`) for the keys listed below; the model is built as
crates/fem/src/grid.rs builds the graded grid (`axis_planes`, `place`, `assign`, `faces`, `crop`,
`locate`), as heat.rs:1483-1597 `build_grid` applies symmetry, heat.rs:1165-1208 `number`
numbers the points, and as heat.rs:328-700 (steady), heat.rs:684-735 (transient: backward Euler
with the lumped capacity, one solver for every step) and elastic.rs:287-563 (static) assemble and
solve. A static step with a temperature (uniform, or the steady heat solve of its heat loads on the
same grid, elastic.rs:358-363) adds each cell's thermal force to the loads (elastic.rs:383-404);
its surfaces give their shape (elastic.rs:1099-1250), its probes their stress (elastic.rs:676-700)
and its displacement loads their reactions (elastic.rs:773-839). The matrices and right sides are assembled
on the device (`fem_assemble`), the solve is the library's (`fem_solver_new_csr` with the grid
level, and the cells for a displacement).

Keys read: `units`; `[grid]` max_cell, max_growth, min_cells, symmetry; `[material.*]`
conductivity (a number; with density and specific_heat for a transient step) or youngs_modulus and
poisson_ratio (1 number each, or 3 each with shear_modulus: orthotropic), expansion (1 or 3) and
conductivity; `[part.*]` material, x, y, z, priority;
`[face.*]` side, box, parts; `[[load]]` temperature, source, flux, convection with sink (heat, and
the temperature of a static step), ux / uy / uz on a face or a point, traction, pressure (static);
`[step]` type "steady", "transient" (end, dt, initial) or "static" (reference_temperature,
temperature); `[probe.*]` point; `[surface.*]` face. Anything else (tables, footprints, radiation,
gravity, Joule heat and other steps) is refused with a message.
"""
from __future__ import annotations

import bisect
import math
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

try:
    import tomllib as _toml  # Python 3.11+
except ImportError:  # pragma: no cover
    import tomli as _toml

VOID = -1
# crates/fem/src/grid.rs:55-81 Side::ALL: x-, x+, y-, y+, z-, z+.
SIDES = [(0, False), (0, True), (1, False), (1, True), (2, False), (2, True)]
SIDE_NAMES = {"x-": 0, "x+": 1, "y-": 2, "y+": 3, "z-": 4, "z+": 5}
VTK = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
STEADY_DIRECT_LIMIT = 3000     # solve.rs:18
TRANSIENT_DIRECT_LIMIT = 40000  # solve.rs:24
DIRECT_FLOPS = 1.5e5           # elastic.rs:263
DIRECT_ENTRIES = 250_000_000   # elastic.rs:269
MULTIGRID_ASPECT = 2.0         # elastic.rs:276


class CaseError(Exception):
    pass


@dataclass
class Part:
    name: str
    material: int
    priority: int
    lo: list
    hi: list


@dataclass
class FaceSet:
    name: str
    sides: list          # indices into SIDES
    region: Optional[tuple]
    parts: list


@dataclass
class Case:
    units: str
    step: str                       # "steady", "transient" or "static"
    max_cell: float
    max_growth: float
    min_cells: int
    symmetry: list
    materials: list                 # dicts: name, and conductivity or youngs_modulus + poisson_ratio
    parts: list
    faces: list
    loads: list                     # dicts: kind and its values
    probes: list                    # (name, point)
    name: str = ""
    transient: Optional[dict] = None  # step "transient": end, dt, initial (heat.rs Step::Transient)
    thermal: Optional[dict] = None    # step "static": {"uniform": T} or {"heat": True}, and "reference"
    surfaces: list = field(default_factory=list)  # (name, face set) of a static step (elastic.rs:93-97)


# ------------------------------------------------------------------------------- reading

def _only(table, where, keys):
    for k in table:
        if k not in keys:
            raise CaseError(f"{where}: `{k}` is not read by this front end (it reads {', '.join(keys)})")


def _number(table, key, where):
    v = table.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise CaseError(f"{where}: `{key}` must be a number")
    return float(v)


def _numbers(v, n, where, key):
    if not isinstance(v, list) or len(v) != n or not all(isinstance(x, (int, float)) and not isinstance(x, bool)
                                                         for x in v):
        raise CaseError(f"{where}: `{key}` must be {n} numbers")
    return [float(x) for x in v]


def _elastic_constants(m, where):
    """core grid.rs:1326-1345: youngs_modulus and poisson_ratio as 1 number each (isotropic), or
    youngs_modulus (x, y, z), poisson_ratio (xy, xz, yz) and shear_modulus (xy, xz, yz) as 3 numbers
    each (orthotropic); the 9 constants E_x, E_y, E_z, nu_xy, nu_xz, nu_yz, G_xy, G_xz, G_yz."""
    def values(key):
        v = m.get(key)
        if v is None:
            return None
        v = v if isinstance(v, list) else [v]
        if not all(isinstance(x, (int, float)) and not isinstance(x, bool) for x in v):
            return []
        return [float(x) for x in v]
    y, p, g = values("youngs_modulus"), values("poisson_ratio"), values("shear_modulus")
    if y is not None and p is not None and g is None and len(y) == 1 and len(p) == 1:
        gi = y[0] / (2.0 * (1.0 + p[0]))
        return [y[0]] * 3 + [p[0]] * 3 + [gi] * 3
    if y is not None and p is not None and g is not None and len(y) == len(p) == len(g) == 3:
        return y + p + g
    raise CaseError(f"{where}: give youngs_modulus and poisson_ratio as 1 number each (isotropic), or "
                    "youngs_modulus (x, y, z), poisson_ratio (xy, xz, yz) and shear_modulus (xy, xz, yz) as 3 "
                    "numbers each (orthotropic)")


def _expansion(m, where):
    """core grid.rs:1347-1355: the expansion along x, y and z, 0 when the material has none."""
    v = m.get("expansion")
    if v is None:
        return [0.0, 0.0, 0.0]
    v = v if isinstance(v, list) else [v]
    if not all(isinstance(x, (int, float)) and not isinstance(x, bool) for x in v) or len(v) not in (1, 3):
        raise CaseError(f"{where}: missing key `expansion`")
    return [float(x) for x in v] * (3 if len(v) == 1 else 1)


def read(text: str, name: str = "") -> Case:
    """The case of a grid.toml text (core/src/grid.rs:79-152 `read`, :153 `read_heat`), for the
    keys this front end reads."""
    top = _toml.loads(text)
    _only(top, "the top", ["units", "grid", "material", "part", "face", "load", "step", "probe", "surface"])
    units = top.get("units")
    if units not in ("SI-m", "SI-mm"):
        raise CaseError(f"`units` must be SI-m or SI-mm, not {units}")
    grid = top.get("grid", {})
    _only(grid, "[grid]", ["max_cell", "max_growth", "min_cells", "symmetry"])
    max_cell = _number(grid, "max_cell", "[grid]")
    max_growth = float(grid["max_growth"]) if "max_growth" in grid else 1.3
    min_cells = grid.get("min_cells", 2)
    if not isinstance(min_cells, int) or min_cells < 1:
        raise CaseError("[grid]: `min_cells` must be a whole number, 1 or more")
    symmetry = [False] * 3
    for a in grid.get("symmetry", []):
        if a not in ("x", "y", "z"):
            raise CaseError("[grid]: `symmetry` must list axes x, y or z")
        symmetry["xyz".index(a)] = True
    step = top.get("step", {})
    kind = step.get("type")
    transient = None
    if kind == "transient":
        _only(step, "[step]", ["type", "end", "dt", "initial", "output_every"])
        transient = {"end": _number(step, "end", "[step]"), "dt": _number(step, "dt", "[step]"),
                     "initial": _number(step, "initial", "[step]")}
        if not (transient["dt"] > 0.0 and transient["end"] > 0.0):
            raise CaseError("[step]: `end` and `dt` must be numbers above 0")
    elif kind == "static":
        _only(step, "[step]", ["type", "reference_temperature", "temperature"])
    else:
        _only(step, "[step]", ["type"])
    if kind not in ("steady", "static", "transient"):
        raise CaseError(f"[step]: type {kind!r}: this front end solves steady and transient heat and static stress")
    materials, mindex = [], {}
    for mname, m in top.get("material", {}).items():
        where = f"[material.{mname}]"
        if kind == "steady":
            _only(m, where, ["conductivity"])
            materials.append({"name": mname, "conductivity": _number(m, "conductivity", where)})
        elif kind == "transient":
            _only(m, where, ["conductivity", "density", "specific_heat"])
            materials.append({"name": mname, "conductivity": _number(m, "conductivity", where),
                              "capacity": _number(m, "density", where) * _number(m, "specific_heat", where)})
        else:
            _only(m, where, ["youngs_modulus", "poisson_ratio", "shear_modulus", "density", "expansion",
                             "conductivity"])
            materials.append({"name": mname, "elastic": _elastic_constants(m, where),
                              "expansion": _expansion(m, where),
                              "conductivity": _number(m, "conductivity", where) if "conductivity" in m else None})
        mindex[mname] = len(materials) - 1
    parts = []
    for pname, p in top.get("part", {}).items():
        where = f"[part.{pname}]"
        _only(p, where, ["material", "x", "y", "z", "priority"])
        if p.get("material") not in mindex:
            raise CaseError(f"{where}: no material {p.get('material')}")
        lo, hi = [0.0] * 3, [0.0] * 3
        for d, ax in enumerate("xyz"):
            r = _numbers(p.get(ax), 2, where, ax)
            lo[d], hi[d] = r
        prio = p.get("priority", 0)
        if not isinstance(prio, int):
            raise CaseError(f"{where}: `priority` must be a whole number")
        parts.append(Part(pname, mindex[p["material"]], prio, lo, hi))
    if not parts:
        raise CaseError("the model has no part")
    pindex = {p.name: i for i, p in enumerate(parts)}
    faces, findex = [], {}
    for fname, f in top.get("face", {}).items():
        where = f"[face.{fname}]"
        _only(f, where, ["side", "box", "parts"])
        sides = f.get("side", [])
        sides = [sides] if isinstance(sides, str) else sides
        for s in sides:
            if s not in SIDE_NAMES:
                raise CaseError(f"{where}: side `{s}` must be x-, x+, y-, y+, z- or z+")
        region = None
        if "box" in f:
            v = _numbers(f["box"], 6, where, "box")
            region = (v[:3], v[3:])
        fparts = []
        for pn in f.get("parts", []):
            if pn not in pindex:
                raise CaseError(f"{where}: no part {pn}")
            fparts.append(pindex[pn])
        faces.append(FaceSet(fname, [SIDE_NAMES[s] for s in sides], region, fparts))
        findex[fname] = len(faces) - 1
    loads = []
    for i, l in enumerate(top.get("load", [])):
        where = f"[[load]] {i + 1}"

        def face_of(l=l, where=where):
            if l.get("face") not in findex:
                raise CaseError(f"{where}: no face set {l.get('face')}")
            return findex[l["face"]]
        if kind in ("steady", "transient"):
            if "temperature" in l:
                _only(l, where, ["temperature", "face"])
                loads.append({"kind": "temperature", "faces": face_of(), "value": _number(l, "temperature", where)})
            elif "source" in l:
                _only(l, where, ["source", "part"])
                if l.get("part") not in pindex:
                    raise CaseError(f"{where}: no part {l.get('part')}")
                loads.append({"kind": "source", "part": pindex[l["part"]], "value": _number(l, "source", where)})
            elif "flux" in l:
                _only(l, where, ["flux", "face"])
                loads.append({"kind": "flux", "faces": face_of(), "value": _number(l, "flux", where)})
            elif "convection" in l:
                _only(l, where, ["convection", "sink", "face"])
                film = _number(l, "convection", where)
                if not film > 0.0:
                    raise CaseError("a convection needs a film coefficient above 0")
                loads.append({"kind": "convection", "faces": face_of(), "film": film, "sink": _number(l, "sink", where)})
            else:
                raise CaseError(f"{where}: a steady load here is temperature, source, flux or convection")
        else:
            if any(k in l for k in ("ux", "uy", "uz")):
                value = [(_number(l, k, where) if k in l else None) for k in ("ux", "uy", "uz")]
                if "point" in l:  # elastic.rs:419-423: a point support
                    _only(l, where, ["ux", "uy", "uz", "point"])
                    loads.append({"kind": "point", "point": _numbers(l["point"], 3, where, "point"), "value": value})
                else:
                    _only(l, where, ["ux", "uy", "uz", "face"])
                    loads.append({"kind": "displacement", "faces": face_of(), "value": value})
            elif any(k in l for k in ("temperature", "source", "flux", "convection")):
                # heat loads of a static step: the temperature of its thermal strain (core grid.rs:814-830)
                if "temperature" in l:
                    _only(l, where, ["temperature", "face"])
                    loads.append({"kind": "temperature", "faces": face_of(), "value": _number(l, "temperature", where),
                                  "heat": True})
                elif "source" in l:
                    _only(l, where, ["source", "part"])
                    if l.get("part") not in pindex:
                        raise CaseError(f"{where}: no part {l.get('part')}")
                    loads.append({"kind": "source", "part": pindex[l["part"]], "value": _number(l, "source", where),
                                  "heat": True})
                elif "flux" in l:
                    _only(l, where, ["flux", "face"])
                    loads.append({"kind": "flux", "faces": face_of(), "value": _number(l, "flux", where), "heat": True})
                else:
                    _only(l, where, ["convection", "sink", "face"])
                    loads.append({"kind": "convection", "faces": face_of(), "film": _number(l, "convection", where),
                                  "sink": _number(l, "sink", where), "heat": True})
            elif "traction" in l:
                _only(l, where, ["traction", "face"])
                loads.append({"kind": "traction", "faces": face_of(), "value": _numbers(l["traction"], 3, where, "traction")})
            elif "pressure" in l:
                _only(l, where, ["pressure", "face"])
                loads.append({"kind": "pressure", "faces": face_of(), "value": _number(l, "pressure", where)})
            else:
                raise CaseError(f"{where}: a static load here is ux / uy / uz, traction or pressure")
    probes = []
    for pname, p in top.get("probe", {}).items():
        where = f"[probe.{pname}]"
        _only(p, where, ["point"])
        probes.append((pname, _numbers(p.get("point"), 3, where, "point")))
    surfaces = []
    for sname, sf in top.get("surface", {}).items():
        where = f"[surface.{sname}]"
        if kind != "static":
            raise CaseError(f"{where}: a surface is a result of a static step: [step] type = \"static\"")
        _only(sf, where, ["face"])
        if sf.get("face") not in findex:
            raise CaseError(f"{where}: no face set {sf.get('face')}")
        surfaces.append((sname, findex[sf["face"]]))
    thermal = None
    if kind == "static":
        heat = [l for l in loads if l.get("heat")]
        uniform = step.get("temperature")
        if uniform is not None and heat:
            raise CaseError("[step]: give `temperature` or heat loads (temperature, flux, convection, emissivity, "
                            "source), not both")
        if uniform is not None or heat:
            if "reference_temperature" not in step:
                raise CaseError("[step]: missing key `reference_temperature`, the temperature with no thermal strain")
            thermal = {"reference": _number(step, "reference_temperature", "[step]")}
            if uniform is not None:
                thermal["uniform"] = _number(step, "temperature", "[step]")
            else:
                thermal["heat"] = True
                if any(m["conductivity"] is None for m in materials):
                    raise CaseError("the heat loads of a static step need the conductivity of every material")
        elif "reference_temperature" in step:
            raise CaseError("[step]: `reference_temperature` needs `temperature` or heat loads")
    return Case(units, kind, max_cell, max_growth, min_cells, symmetry, materials, parts, faces, loads, probes, name,
                transient, thermal, surfaces)


def load_case(path: str) -> Case:
    """The case of a grid.toml file, named after its folder."""
    with open(path, "rb") as f:
        text = f.read().decode()
    return read(text, os.path.basename(os.path.dirname(os.path.abspath(path))))


# ------------------------------------------------------------------------------- the grid

def tol(p) -> float:
    """grid.rs:405-408: 1e-9 of the length of the planes."""
    return 1e-9 * max(abs(p[-1] - p[0]), 2.2250738585072014e-308)


def place(keys, target, at_key, slope):
    """grid.rs:492-521: the planes of an axis from the size function h."""
    def h(x, i):
        v = target[i]
        for k, s in zip(keys, at_key):
            v = min(v, s + slope * abs(x - k))
        return v
    planes = [keys[0]]
    for i in range(len(keys) - 1):
        a, b = keys[i], keys[i + 1]
        samples = 400
        step = (b - a) / float(samples)
        phi = [0.0] * (samples + 1)
        for s in range(samples):
            x0, x1 = a + s * step, a + (s + 1) * step
            phi[s + 1] = phi[s] + 0.5 * step * (1.0 / h(x0, i) + 1.0 / h(x1, i))
        total = phi[samples]
        cells = max(int(math.ceil(total - 1e-6)), 1)
        for c in range(1, cells):
            want = total * c / cells
            s = min(max(bisect.bisect_left(phi, want), 1), samples)
            f = (want - phi[s - 1]) / (phi[s] - phi[s - 1])
            planes.append(a + (s - 1.0 + f) * step)
        planes.append(b)
    return planes


def axis_planes(parts, d, max_cell, max_growth, min_cells, extra):
    """grid.rs:411-489: the graded planes of axis d."""
    keys = sorted([v for p in parts for v in (p.lo[d], p.hi[d])] + list(extra))
    t = tol(keys)
    dedup = [keys[0]]
    for v in keys[1:]:
        if abs(v - dedup[-1]) > t:
            dedup.append(v)
    keys = dedup
    target = []
    for w0, w1 in zip(keys[:-1], keys[1:]):
        v = max_cell
        for p in parts:
            if p.lo[d] <= w0 + t and p.hi[d] >= w1 - t:
                v = min(v, (p.hi[d] - p.lo[d]) / float(min_cells))
        target.append(v)
    at_key = []
    for k in range(len(keys)):
        left = target[k - 1] if k > 0 else math.inf
        right = target[k] if k < len(target) else math.inf
        at_key.append(min(left, right))
    slope = 0.5 * (max_growth - 1.0)
    for _ in range(60):
        planes = place(keys, target, at_key, slope)
        cells = [b - a for a, b in zip(planes[:-1], planes[1:])]
        fixed = False
        for c in range(len(cells) - 1):
            small, large = (cells[c], cells[c + 1]) if cells[c] < cells[c + 1] else (cells[c + 1], cells[c])
            if large > max_growth * small * (1.0 + 1e-9):
                x = planes[c + 1]
                k = min(bisect.bisect_left(keys, x - t), len(keys) - 1)
                for kk in (max(k - 1, 0), k, min(k + 1, len(keys) - 1)):
                    if abs(keys[kk] - x) <= 2.0 * large and at_key[kk] > small * 1.0001:
                        at_key[kk] = small
                        fixed = True
        if not fixed:
            return planes
    raise CaseError(f"the {'xyz'[d]} grid planes break the growth limit max_growth = {max_growth}: use a larger "
                    "max_growth or a smaller max_cell")


class Grid:
    """grid.rs:108-403: the planes of each axis and the part of each cell (x fastest) or VOID."""

    def __init__(self, planes, part):
        self.planes = [list(p) for p in planes]
        self.part = np.asarray(part, np.int64)

    @classmethod
    def build(cls, case: Case, extra):
        planes = [axis_planes(case.parts, d, case.max_cell, case.max_growth, case.min_cells, extra[d]) for d in range(3)]
        g = cls(planes, np.full((len(planes[0]) - 1) * (len(planes[1]) - 1) * (len(planes[2]) - 1), VOID))
        g._assign(case.parts)
        return g

    def dims(self):
        return [len(p) - 1 for p in self.planes]

    def cell_index(self, c):
        nx, ny, _ = self.dims()
        return c[0] + nx * (c[1] + ny * c[2])

    def cell_ijk(self, cell):
        nx, ny, _ = self.dims()
        return [cell % nx, (cell // nx) % ny, cell // (nx * ny)]

    def point_index(self, p):
        nx, ny, _ = self.dims()
        return p[0] + (nx + 1) * (p[1] + (ny + 1) * p[2])

    def size(self, cell):
        c = self.cell_ijk(cell)
        return [self.planes[d][c[d] + 1] - self.planes[d][c[d]] for d in range(3)]

    def centre(self, cell):
        c = self.cell_ijk(cell)
        return [0.5 * (self.planes[d][c[d]] + self.planes[d][c[d] + 1]) for d in range(3)]

    def corners(self, cell):
        i, j, k = self.cell_ijk(cell)
        return [(i + b[0], j + b[1], k + b[2]) for b in VTK]

    def _assign(self, parts):
        """grid.rs:290-327: each cell takes the part that holds its centre (priority, then order)."""
        best = [None] * len(self.part)
        for index, p in enumerate(parts):
            rng = []
            for d in range(3):
                pl = self.planes[d]
                t = tol(pl)
                rng.append((bisect.bisect_left(pl, p.lo[d] - t), bisect.bisect_left(pl, p.hi[d] - t)))
            for k in range(*rng[2]):
                for j in range(*rng[1]):
                    for i in range(*rng[0]):
                        cell = self.cell_index([i, j, k])
                        other = best[cell]
                        if other is not None and parts[other].priority > p.priority:
                            continue
                        if other is not None and parts[other].priority == p.priority and \
                                parts[other].material != p.material:
                            x = self.centre(cell)
                            raise CaseError(f"parts {parts[other].name} and {p.name} overlap at ({x[0]}, {x[1]}, {x[2]}) "
                                            "with the same priority and other materials: give the part that must "
                                            "hold the overlap a higher priority")
                        best[cell] = index
        self.part = np.array([VOID if b is None else b for b in best], np.int64)

    def neighbour(self, cell, side):
        axis, positive = SIDES[side]
        c = self.cell_ijk(cell)
        n = self.dims()[axis]
        if positive:
            if c[axis] + 1 >= n:
                return None
            c[axis] += 1
        else:
            if c[axis] == 0:
                return None
            c[axis] -= 1
        return self.cell_index(c)

    def faces(self):
        """grid.rs:233-253: each side of a cell with a part and no part beyond it, cells in order."""
        out = []
        for cell in range(len(self.part)):
            if self.part[cell] == VOID:
                continue
            for s in range(6):
                n = self.neighbour(cell, s)
                if n is None or self.part[n] == VOID:
                    out.append((cell, s))
        return out

    def face_centre(self, face):
        cell, s = face
        axis, positive = SIDES[s]
        c = self.cell_ijk(cell)
        x = self.centre(cell)
        x[axis] = self.planes[axis][c[axis] + (1 if positive else 0)]
        return x

    def face_area(self, face):
        s = self.size(face[0])
        axis = SIDES[face[1]][0]
        v = 1.0
        for d in range(3):
            if d != axis:
                v *= s[d]
        return v

    def face_points(self, face):
        """grid.rs:272-289: (0, 0), (1, 0), (1, 1), (0, 1) along the face's two axes."""
        cell, s = face
        axis, positive = SIDES[s]
        a, b = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
        base = self.cell_ijk(cell)
        base[axis] += 1 if positive else 0
        out = []
        for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
            p = list(base)
            p[a] += da
            p[b] += db
            out.append(tuple(p))
        return out

    def select(self, all_faces, fs: FaceSet):
        """heat.rs:1575-1604 `select`."""
        out = []
        for f in all_faces:
            if fs.sides and f[1] not in fs.sides:
                continue
            if fs.parts and int(self.part[f[0]]) not in fs.parts:
                continue
            if fs.region is not None:
                c = self.face_centre(f)
                if any(c[d] < fs.region[0][d] - tol(self.planes[d]) or c[d] > fs.region[1][d] + tol(self.planes[d])
                       for d in range(3)):
                    continue
            out.append(f)
        return out

    def planes_mirror(self, axis, at):
        p = self.planes[axis]
        n = len(p)
        return all(abs((p[i] - at) + (p[n - 1 - i] - at)) <= tol(p) for i in range(n))

    def mirror_cell(self, cell, axis):
        c = self.cell_ijk(cell)
        c[axis] = self.dims()[axis] - 1 - c[axis]
        return self.cell_index(c)

    def crop(self, axis, at):
        """grid.rs:329-357: only the cells above the plane `at` stay."""
        p = self.planes[axis]
        first = next(i for i, x in enumerate(p) if abs(x - at) <= tol(p))
        planes = [list(q) for q in self.planes]
        planes[axis] = p[first:]
        nx, ny, nz = self.dims()
        part, old = [], []
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    if (i, j, k)[axis] >= first:
                        cell = self.cell_index([i, j, k])
                        part.append(int(self.part[cell]))
                        old.append(cell)
        return Grid(planes, part), old

    def locate(self, x):
        """grid.rs:361-402: the cell of a part that holds x, and x's local coordinates (0 to 1)."""
        choices = []
        for d in range(3):
            p = self.planes[d]
            t = tol(p)
            if x[d] < p[0] - t or x[d] > p[-1] + t:
                return None
            i = min(max(bisect.bisect_right(p, x[d]), 1), len(p) - 1) - 1
            ch = [i]
            if i > 0 and abs(x[d] - p[i]) <= t:
                ch.append(i - 1)
            if i + 2 < len(p) and abs(x[d] - p[i + 1]) <= t:
                ch.append(i + 1)
            choices.append(ch)
        for i in choices[0]:
            for j in choices[1]:
                for k in choices[2]:
                    cell = self.cell_index([i, j, k])
                    if self.part[cell] != VOID:
                        c = (i, j, k)
                        local = []
                        for d in range(3):
                            a, b = self.planes[d][c[d]], self.planes[d][c[d] + 1]
                            local.append(min(max((x[d] - a) / (b - a), 0.0), 1.0))
                        return cell, local
        return None


def build_grid(case: Case, used, keys):
    """heat.rs:1483-1573: the grid with the mid planes of symmetry, the faces of each face set,
    and the grid cut at each plane of symmetry (the part above it stays)."""
    mids = [None] * 3
    for d in range(3):
        if case.symmetry[d]:
            lo = min(p.lo[d] for p in case.parts)
            hi = max(p.hi[d] for p in case.parts)
            mids[d] = 0.5 * (lo + hi)
    extra = [[m] if m is not None else [] for m in mids]
    grid = Grid.build(case, extra)
    all_faces = grid.faces()
    faces = [grid.select(all_faces, fs) for fs in case.faces]
    for fs, lst, u in zip(case.faces, faces, used):
        if u and not lst:
            raise CaseError(f"face set {fs.name} selects no face of the model: a face of the model has no part beyond it")
    for d in range(3):
        mid = mids[d]
        if mid is None:
            continue
        axis = "xyz"[d]
        if not grid.planes_mirror(d, mid):
            raise CaseError(f"symmetry at {axis} = {mid}: the grid planes do not mirror at this plane. Solve the full model")

        def sign(c):
            p = int(grid.part[c])
            return None if p == VOID else (case.parts[p].material, keys[p])
        if any(sign(c) != sign(grid.mirror_cell(c, d)) for c in range(len(grid.part))):
            raise CaseError(f"symmetry at {axis} = {mid}: the materials or the loads of the parts do not mirror at "
                            "this plane. Solve the full model")
        for fs, lst in zip(case.faces, faces):
            have = set(lst)
            for cell, s in lst:
                ax, pos = SIDES[s]
                ms = SIDES.index((ax, not pos)) if ax == d else s
                if (grid.mirror_cell(cell, d), ms) not in have:
                    raise CaseError(f"symmetry at {axis} = {mid}: the faces of face set {fs.name} do not mirror at "
                                    "this plane. Solve the full model")
        cut, old = grid.crop(d, mid)
        new = {o: k for k, o in enumerate(old)}
        faces = [[(new[c], s) for c, s in lst if c in new] for lst in faces]
        grid = cut
    return grid, faces, mids


def number(grid: Grid):
    """heat.rs:1165-1208: points first seen over the cells with a part, corners in VTK order."""
    number_of = {}
    ijk, cells, cell_of = [], [], []
    for cell in range(len(grid.part)):
        if grid.part[cell] == VOID:
            continue
        for p in grid.corners(cell):
            slot = number_of.get(p)
            if slot is None:
                slot = len(ijk)
                number_of[p] = slot
                ijk.append(p)
            cells.append(slot)
        cell_of.append(cell)
    return np.array(ijk, np.int64).reshape(-1, 3), np.array(cells, np.int64).reshape(-1, 8), cell_of, number_of


def probe_weights(case: Case, grid: Grid, mids, number_of):
    """heat.rs:1312-1358: for each probe its (node, weight) pairs at the cell's corners (the
    trilinear functions at its local point), the axes on which it reads its mirror point, its cell
    and its local point (each coordinate -1 to 1)."""
    out = []
    for name, point in case.probes:
        x = list(point)
        mirrored = [False] * 3
        for d in range(3):
            if mids[d] is not None and x[d] < mids[d]:
                x[d] = 2.0 * mids[d] - x[d]
                mirrored[d] = True
        hit = grid.locate(x)
        if hit is None:
            raise CaseError(f"probe {name} at ({point[0]}, {point[1]}, {point[2]}) is in no part")
        cell, local = hit
        xi = [2.0 * f - 1.0 for f in local]
        w = []
        for b in VTK:
            v = 1.0
            for d in range(3):
                v *= 0.5 * (1.0 + xi[d]) if b[d] == 1 else 0.5 * (1.0 - xi[d])
            w.append(v)
        out.append((name, [(number_of[c], wc) for c, wc in zip(grid.corners(cell), w)], mirrored, cell, xi))
    return out


# ------------------------------------------------------------------------------- the solves

@dataclass
class Result:
    """The probes of a solve (column name -> value), the field at the points, and the solve."""
    probes: dict
    field: np.ndarray
    ijk: np.ndarray
    grid: Grid
    method: str
    stats: object = None
    unknowns: int = 0
    notes: list = field(default_factory=list)


def _face_mass(a, b):
    """heat.rs:1280-1292 face_mass(a, b, 1): the 4 x 4 mass of a face in face_points order."""
    bits = ((0, 0), (1, 0), (1, 1), (0, 1))
    ms = lambda l, p, q: l / 3.0 if p == q else l / 6.0
    return [[1.0 * ms(a, bp[0], bq[0]) * ms(b, bp[1], bq[1]) for bq in bits] for bp in bits]


def solve(case: Case, precision: str = "fp32", method: Optional[str] = None, tol: float = 1e-10,
          device: int = 0) -> Result:
    """The case solved on the device: steady heat (heat.rs:328-700) or static stress
    (elastic.rs:287-563)."""
    if case.step in ("steady", "transient"):
        return _solve_heat(case, precision, method, tol, device)
    return _solve_elastic(case, precision, method, tol, device)


def _entry(ptr, idx, r, c):
    lo, hi = int(ptr[r]), int(ptr[r + 1])
    k = lo + int(np.searchsorted(idx[lo:hi], c))
    assert k < hi and idx[k] == c
    return k


def _solve_heat(case, precision, method, tol_, device):
    import femahoy_amd as fem
    used = [any(l.get("faces") == f for l in case.loads) for f in range(len(case.faces))]
    part_source = [0.0] * len(case.parts)
    for l in case.loads:
        if l["kind"] == "source":
            part_source[l["part"]] += l["value"]
    keys = [(q,) for q in part_source]
    grid, faces, mids = build_grid(case, used, keys)
    ijk, cells, cell_of, number_of = number(grid)
    n = len(ijk)
    xyz = np.array([[grid.planes[d][p[d]] for d in range(3)] for p in ijk])
    kc = np.array([case.materials[case.parts[int(grid.part[c])].material]["conductivity"] for c in cell_of])
    # heat.rs:384-426: the volume sources, cell by cell, V/8 to each corner
    source = np.zeros(n)
    for e, c in enumerate(cell_of):
        s = grid.size(c)
        vol = s[0] * s[1] * s[2]
        q = part_source[int(grid.part[c])]
        for node in cells[e]:
            source[node] += q * vol / 8.0
    flux = np.zeros(n)
    film = np.zeros(n)
    fixed = [None] * n
    convection = []
    for l in case.loads:
        if l["kind"] == "source":
            continue
        for face in faces[l["faces"]]:
            nodes = [number_of[p] for p in grid.face_points(face)]
            share = grid.face_area(face) / 4.0
            if l["kind"] == "temperature":
                for v in nodes:
                    if fixed[v] is not None and fixed[v][0] != l["value"]:
                        raise CaseError(f"the face sets {case.faces[fixed[v][1]].name} and {case.faces[l['faces']].name} "
                                        f"share a point with 2 fixed temperatures, {fixed[v][0]} and {l['value']}")
                    fixed[v] = (l["value"], l["faces"])
            elif l["kind"] == "flux":
                for v in nodes:
                    flux[v] += l["value"] * share
            else:
                s = grid.size(face[0])
                axis = SIDES[face[1]][0]
                a, b = {0: (s[1], s[2]), 1: (s[0], s[2]), 2: (s[0], s[1])}[axis]
                convection.append((nodes, _face_mass(a, b), l["film"]))
                for v in nodes:
                    film[v] += l["film"] * l["sink"] * share
    load = source + flux + film
    isfixed = np.array([f is not None for f in fixed], np.uint8)
    value = np.array([f[0] if f is not None else 0.0 for f in fixed])
    mesh = fem.Mesh("heat_hex8", xyz, cells, kc, isfixed, value, load=load)
    asm = fem.Assembly(mesh, device)
    ptr, idx, val, rhs, free = asm.arrays()
    renum = -np.ones(n, np.int64)
    renum[free] = np.arange(len(free))
    # heat.rs:465-473: the convection matrices h m added after the conduction, face by face; the
    # held points' share moves to the right side
    for nodes, m, h in convection:
        for p in range(4):
            rp = renum[nodes[p]]
            if rp < 0:
                continue
            for q in range(4):
                rq = renum[nodes[q]]
                hm = h * m[p][q]
                if rq >= 0:
                    val[_entry(ptr, idx, rp, rq)] += hm
                else:
                    rhs[rp] -= hm * value[nodes[q]]
    planes = [np.asarray(p, np.float64) for p in grid.planes]
    g = fem.Grid(planes, ijk[free].astype(np.uint32), np.zeros(len(free), np.uint8))
    if case.transient is None:
        s = fem.CsrSolver(ptr, idx, val, "grid", g, method=method, precision=precision, tol=tol_,
                          crossover=STEADY_DIRECT_LIMIT, device=device)
        x, st = s.solve(rhs)
        t = value.copy()
        t[free] = x
    else:
        # heat.rs:684-735: the lumped capacity (rho c V / 8 to each corner, cell by cell) over dt on
        # the diagonal, one solver for every step, each step's right side the load, the capacity
        # term of the last temperatures and the fixed points' couplings, its start the last solution
        dt, steps = case.transient["dt"], int(round(case.transient["end"] / case.transient["dt"]))
        capacity = np.zeros(n)
        for e, c in enumerate(cell_of):
            sz = grid.size(c)
            vol = sz[0] * sz[1] * sz[2]
            q = case.materials[case.parts[int(grid.part[c])].material]["capacity"]
            for node in cells[e]:
                capacity[node] += q * vol / 8.0
        for f_, i in enumerate(free):
            val[_entry(ptr, idx, f_, f_)] += capacity[i] / dt
        s = fem.CsrSolver(ptr, idx, val, "grid", g, method=method, precision=precision, tol=tol_,
                          crossover=TRANSIENT_DIRECT_LIMIT, device=device)
        t = np.where(isfixed.astype(bool), value, case.transient["initial"])
        x = t[free].copy()
        cf = capacity[free] / dt
        for _ in range(steps):
            x, st = s.solve(rhs + cf * t[free], x=x)
            t[free] = x
    probes = {}
    for name, pairs, *_ in probe_weights(case, grid, mids, number_of):
        probes[name] = sum(t[v] * w for v, w in pairs)
    return Result(probes, t, ijk, grid, s.method(), st, len(free))


def grid_point(grid: Grid, number_of, mids, x):
    """elastic.rs:1410-1444: the point of the grid at a point support (its mirror point above each
    plane of symmetry), the nearest plane on each axis within 1e-9 of the model's span."""
    y = list(x)
    for d in range(3):
        if mids[d] is not None and y[d] < mids[d]:
            y[d] = 2.0 * mids[d] - y[d]
    span = max(grid.planes[d][-1] - grid.planes[d][0] for d in range(3))
    near = tuple(min(range(len(grid.planes[d])), key=lambda a: abs(grid.planes[d][a] - y[d])) for d in range(3))
    node = number_of.get(near)
    if node is None or any(abs(grid.planes[d][near[d]] - y[d]) > 1e-9 * span for d in range(3)):
        raise CaseError(f"the point support at ({x[0]:e}, {x[1]:e}, {x[2]:e}) is not a point of the grid of the "
                        "parts. A point of the grid is a corner of a part, or a point where the planes of 2 parts "
                        "cross")
    return node


def _thermal_change(case, n, precision, tol_, device):
    """elastic.rs:358-363, 869-906: the temperature change from the reference temperature at each
    point: a uniform temperature, or the steady heat solve of the heat loads on the same parts
    (core grid.rs:816-830), or none."""
    if case.thermal is None:
        return None
    if "uniform" in case.thermal:
        return np.full(n, case.thermal["uniform"] - case.thermal["reference"])
    import dataclasses
    heat = dataclasses.replace(case, step="steady", loads=[l for l in case.loads if l.get("heat")], probes=[],
                               thermal=None, transient=None)
    t = _solve_heat(heat, precision, None, tol_, device).field
    if len(t) != n:
        raise CaseError("the heat model has other parts or another mesh")
    return t - case.thermal["reference"]


def _solve_small(m, r, k):
    """elastic.rs:1252-1286: the first k rows and columns of m x = r by Gauss-Jordan elimination
    with partial pivoting; a column with no pivot above 1e-12 of the largest diagonal gives 0."""
    a = [list(m[i][:k]) + [0.0] * (6 - k) + [r[i]] for i in range(k)]
    scale = max(abs(m[i][i]) for i in range(k))
    x = [0.0] * 6
    pivot_of = [None] * 6
    row = 0
    for col in range(k):
        if row >= k:
            break
        p = max(range(row, k), key=lambda i: (abs(a[i][col]), i))
        if abs(a[p][col]) <= 1e-12 * scale:
            continue
        a[row], a[p] = a[p], a[row]
        for i in range(k):
            if i != row:
                f = a[i][col] / a[row][col]
                for j in range(7):
                    a[i][j] -= f * a[row][j]
        pivot_of[col] = row
        row += 1
    for col in range(k):
        if pivot_of[col] is not None:
            x[col] = a[pivot_of[col]][6] / a[pivot_of[col]][col]
    return x


def surface_shape(grid: Grid, faces, mids, number_of, u, name):
    """elastic.rs:1099-1250: the shape of a surface from w, the displacement along its outward
    normal: its faces and their mirrors across the planes of symmetry in its plane, the least
    squares plane and quadratic of w on each face's 2 x 2 Gauss points, and the range of w and of w
    minus the plane at the corners. The columns of surfaces.txt (fh grid.rs:846-862)."""
    side = faces[0][1]
    if any(f[1] != side for f in faces):
        raise CaseError(f"surface {name}: its faces must all face 1 direction, for example z+: give its face set "
                        "a `side`")
    axis, positive = SIDES[side]
    sign = 1.0 if positive else -1.0
    ia, ib = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
    quads = []
    for face in faces:
        q = []
        for p in grid.face_points(face):
            x = [grid.planes[d][p[d]] for d in range(3)]
            q.append([x[ia], x[ib], sign * u[3 * number_of[p] + axis]])
        quads.append(q)
    for k, d in enumerate((ia, ib)):
        if mids[d] is not None:
            m = mids[d]
            copies = []
            for q in quads:
                c = [list(v) for v in q]
                for v in c:
                    v[k] = 2.0 * m - v[k]
                copies.append(c)
            quads.extend(copies)
    g = 0.5 / math.sqrt(3.0)
    samples = []
    for q in quads:
        da = q[1][0] - q[0][0]
        db = q[3][1] - q[0][1]
        weight = abs(da * db) / 4.0
        for fa, fb in ((0.5 - g, 0.5 - g), (0.5 + g, 0.5 - g), (0.5 + g, 0.5 + g), (0.5 - g, 0.5 + g)):
            nsh = ((1.0 - fa) * (1.0 - fb), fa * (1.0 - fb), fa * fb, (1.0 - fa) * fb)
            w = 0.0
            for k in range(4):
                w += nsh[k] * q[k][2]
            samples.append((q[0][0] + fa * da, q[0][1] + fb * db, w, weight))
    nodes = [v for q in quads for v in q]
    area = ca = cb = 0.0
    for a_, b_, _, wt in samples:
        area += wt
        ca += wt * a_
        cb += wt * b_
    ca, cb = ca / area, cb / area
    lo = [min(c[k] for c in nodes) for k in range(2)]
    hi = [max(c[k] for c in nodes) for k in range(2)]
    half = max(0.5 * max(hi[0] - lo[0], hi[1] - lo[1]), 2.2250738585072014e-308)

    def basis(a_, b_):
        s_, t_ = (a_ - ca) / half, (b_ - cb) / half
        return (1.0, s_, t_, s_ * s_, s_ * t_, t_ * t_)
    mm = [[0.0] * 6 for _ in range(6)]
    rr = [0.0] * 6
    for a_, b_, w, wt in samples:
        phi = basis(a_, b_)
        for i in range(6):
            rr[i] += wt * phi[i] * w
            for j in range(6):
                mm[i][j] += wt * phi[i] * phi[j]
    plane = _solve_small(mm, rr, 3)
    quad = _solve_small(mm, rr, 6)
    wmin = rmin = math.inf
    wmax = rmax = -math.inf
    for c in nodes:
        phi = basis(c[0], c[1])
        fit = 0.0
        for k in range(3):
            fit += plane[k] * phi[k]
        wmin, wmax = min(wmin, c[2]), max(wmax, c[2])
        rmin, rmax = min(rmin, c[2] - fit), max(rmax, c[2] - fit)
    h2 = half * half
    curv = (-2.0 * quad[3] / h2, -2.0 * quad[5] / h2, -quad[4] / h2)
    mean = 0.5 * (curv[0] + curv[1])
    radius = math.sqrt(0.25 * (curv[0] - curv[1]) ** 2 + curv[2] * curv[2])
    a_n, b_n = "xyz"[ia], "xyz"[ib]
    k1, k2 = mean + radius, mean - radius
    return {f"{name}.range": wmax - wmin, f"{name}.flatness": rmax - rmin, f"{name}.k{a_n}{a_n}": curv[0],
            f"{name}.k{b_n}{b_n}": curv[1], f"{name}.k{a_n}{b_n}": curv[2], f"{name}.k1": k1, f"{name}.k2": k2,
            f"{name}.shape": "convex" if k1 + k2 > 0.0 else ("concave" if k1 + k2 < 0.0 else "flat")}


def _solve_elastic(case, precision, method, tol_, device):
    import femahoy_amd as fem
    used = [any(l.get("faces") == f for l in case.loads) or any(sf == f for _, sf in case.surfaces)
            for f in range(len(case.faces))]
    keys = [(0.0, 0.0, 0.0) for _ in case.parts]
    grid, faces, mids = build_grid(case, used, keys)
    ijk, cells, cell_of, number_of = number(grid)
    n = len(ijk)
    xyz = np.array([[grid.planes[d][p[d]] for d in range(3)] for p in ijk])
    mat = [case.materials[case.parts[int(grid.part[c])].material] for c in cell_of]
    ortho = np.array([m["elastic"] for m in mat])
    # each distinct (cell size, material) once: the element and its thermal load
    table = {}
    which = []
    for c in cell_of:
        key = (tuple(grid.size(c)), case.parts[int(grid.part[c])].material)
        if key not in table:
            m = case.materials[key[1]]
            el = m["elastic"]
            table[key] = fem.hex_element(key[0], el[0:3], el[3:6], m["expansion"], g=el[6:9])
        which.append(key)
    dt = _thermal_change(case, n, precision, tol_, device)
    force = np.zeros(3 * n)
    if dt is not None:
        # elastic.rs:383-404: each cell's thermal force, load x the temperature change at its corners,
        # in cell order before the surface loads
        for e, key in enumerate(which):
            f = table[key][1] @ dt[cells[e]]
            for r in range(24):
                force[3 * cells[e][r // 3] + r % 3] += f[r]
    fixed = [None] * (3 * n)
    owner = [-1] * (3 * n)  # the load that first fixes each unknown, for the reactions (elastic.rs:419)

    def fix(v, value, li):
        for d in range(3):
            val = value[d]
            if val is None:
                continue
            if fixed[3 * v + d] is not None and fixed[3 * v + d] != val:
                x = xyz[v]
                raise CaseError(f"2 fixed displacements give 2 values of u{'xyz'[d]} at the point "
                                f"({x[0]:e}, {x[1]:e}, {x[2]:e}): {fixed[3 * v + d]} and {val}")
            if fixed[3 * v + d] is None:
                owner[3 * v + d] = li
            fixed[3 * v + d] = val
    for li, l in enumerate(case.loads):
        if l.get("heat"):
            continue
        if l["kind"] == "point":
            fix(grid_point(grid, number_of, mids, l["point"]), l["value"], li)
            continue
        for face in faces[l["faces"]]:
            pts = grid.face_points(face)
            if l["kind"] == "displacement":
                for p in pts:
                    fix(number_of[p], l["value"], li)
            else:
                if l["kind"] == "traction":
                    t = list(l["value"])
                else:  # elastic.rs:448-459: a pressure pushes along the inner normal
                    axis, positive = SIDES[face[1]]
                    t = [0.0, 0.0, 0.0]
                    t[axis] = -l["value"] if positive else l["value"]
                area = grid.face_area(face)
                for p in pts:
                    v = number_of[p]
                    for d in range(3):
                        force[3 * v + d] += t[d] * area / 4.0
    # elastic.rs:493-511: the planes of symmetry hold the normal component
    for i in range(n):
        for d in range(3):
            if mids[d] is not None and ijk[i][d] == 0:
                f = fixed[3 * i + d]
                if f is not None and f != 0.0:
                    raise CaseError(f"a fixed u{'xyz'[d]} = {f} is on the plane of symmetry: it does not mirror. "
                                    "Solve the full model")
                fixed[3 * i + d] = 0.0
    isfixed = np.array([f is not None for f in fixed], np.uint8)
    value = np.array([f if f is not None else 0.0 for f in fixed])
    mesh = fem.Mesh("elastic_hex8", xyz, cells, None, isfixed, value, load=force, ortho=ortho)
    asm = fem.Assembly(mesh, device)
    ptr, idx, val, rhs, free = asm.arrays()
    planes = [np.asarray(p, np.float64) for p in grid.planes]
    g = fem.Grid(planes, ijk[free // 3].astype(np.uint32), (free % 3).astype(np.uint8))
    # elastic.rs:525-547: Crossover::Flops, flat when a cell's sides differ by more than 2 times
    aspect = 1.0
    for c in cell_of:
        s = grid.size(c)
        aspect = max(aspect, max(s) / min(s))
    origin = np.array([grid.cell_ijk(c) for c in cell_of], np.uint32)
    element = np.stack([table[key][0] for key in which])
    crossover = fem.Flops(flops=DIRECT_FLOPS, entries=DIRECT_ENTRIES, flat=aspect > MULTIGRID_ASPECT)
    s = fem.CsrSolver(ptr, idx, val, "grid", g, method=method, precision=precision, tol=tol_, crossover=crossover,
                      device=device, cells=(origin, element))
    x, st = s.solve(rhs)
    u = value.copy()
    u[free] = x
    probes = {}
    for name, pairs, mirrored, cell, xi in probe_weights(case, grid, mids, number_of):
        for d, comp in enumerate(("ux", "uy", "uz")):
            v = sum(u[3 * node + d] * w for node, w in pairs)
            probes[f"{name}.{comp}"] = -v if mirrored[d] else v
        # elastic.rs:676-700: the stress of the probe's cell at its local point, in the order xx, yy,
        # zz, xy, yz, xz (elastic.rs:967); a mirrored axis turns the shears that hold it
        nodes = [node for node, _ in pairs]
        ue = np.array([u[3 * node + c] for node in nodes for c in range(3)])
        te = dt[nodes] if dt is not None else np.zeros(8)
        m = case.materials[case.parts[int(grid.part[cell])].material]
        el = m["elastic"]
        _, sg = fem.hex_strain_stress(grid.size(cell), el[0:3], el[3:6], m["expansion"], xi, ue, te, g=el[6:9])
        sig = [sg[0], sg[1], sg[2], sg[5], sg[3], sg[4]]
        for a in range(3):
            if mirrored[a]:
                for sidx, pair in enumerate(((0, 1), (1, 2), (0, 2))):
                    if a in pair:
                        sig[3 + sidx] = -sig[3 + sidx]
        for comp, v in zip(("sxx", "syy", "szz", "sxy", "syz", "sxz"), sig):
            probes[f"{name}.{comp}"] = v
        xx, yy, zz, xy, yz, xz = sig
        probes[f"{name}.mises"] = math.sqrt(0.5 * ((xx - yy) ** 2 + (yy - zz) ** 2 + (zz - xx) ** 2)
                                            + 3.0 * (xy * xy + yz * yz + xz * xz))
        if dt is not None:
            probes[f"{name}.t"] = sum(w * (dt[node] + case.thermal["reference"]) for node, w in pairs)
    # elastic.rs:773-839: the reactions K u - f at the unknowns that each displacement load fixes
    # first (planes of symmetry and point supports have none), named after its face set
    dofs = (3 * cells[:, :, None] + np.arange(3)).reshape(len(cells), 24)
    ku = np.zeros(3 * n)
    np.add.at(ku, dofs, np.einsum("eij,ej->ei", element, u[dofs]))
    for li, l in enumerate(case.loads):
        if l["kind"] != "displacement":
            continue
        f = [0.0, 0.0, 0.0]
        for i in range(3 * n):
            if owner[i] == li:
                f[i % 3] += ku[i] - force[i]
        name = case.faces[l["faces"]].name
        for comp, v in zip(("fx", "fy", "fz"), f):
            probes[f"{name}.{comp}"] = v
    for name, sf in case.surfaces:
        probes.update(surface_shape(grid, faces[sf], mids, number_of, u, name))
    return Result(probes, u, ijk, grid, s.method(), st, len(free))


def check(result: Result, check_text: str, case_name: str):
    """The entries of a check.toml (crates/fh/src/validation/run.rs:38: pass when |computed -
    reference| <= rel_tol |reference| + abs_tol) that name this case and a column of its probes:
    a list of (entry, column, computed, reference, limit, passed)."""
    doc = _toml.loads(check_text)
    rel, ab = float(doc.get("rel_tol", 0.0)), float(doc.get("abs_tol", 0.0))
    out = []
    for key, item in doc.items():
        if not isinstance(item, dict) or item.get("case") != case_name or item.get("column") not in result.probes:
            continue
        got = result.probes[item["column"]]
        ref = float(item["value"])
        limit = rel * abs(ref) + ab
        out.append((key, item["column"], got, ref, limit, abs(got - ref) <= limit))
    return out
