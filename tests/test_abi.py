"""The C ABI: the built library loads and exports every function the public headers declare,
the Python mirror binds them all, the defaults are the reference's constants, and the product
library neither links nor names the oracle. CPU only (no device call)."""
import ctypes as C
import os
import re
import subprocess

import pytest


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("fem_solver.h", "fem_testing.h")]


def declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported(fem):
    lib = C.CDLL(fem.LIB_PATH)
    names = [n for h in HEADERS for n in declared(h)]
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_the_python_mirror_binds_the_public_header(fem):
    public = set(declared(HEADERS[0]))
    assert public == set(fem.EXPORTS)
    for n in fem.EXPORTS:
        assert getattr(fem.lib(), n).argtypes is not None, n


def test_the_defaults_are_the_reference_constants(fem):
    o = fem._Options()
    fem.lib().fem_options_default(C.byref(o))
    assert o.tol == 1e-10          # solve.rs:42 PCG_TOL
    assert o.max_iter == 500       # solve.rs:43 PCG_MAX_ITER
    assert o.crossover == 3000     # solve.rs:18 STEADY_DIRECT_LIMIT
    assert o.method == -1
    assert (fem.PCG_TOL, fem.PCG_MAX_ITER, fem.STEADY_DIRECT_LIMIT) == (1e-10, 500, 3000)


def test_the_library_does_not_link_the_oracle(fem):
    out = subprocess.run(["ldd", fem.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", "--defined-only", fem.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms
    assert re.search(r"\bfem_solver_new_box\b", syms)


def test_the_library_holds_gfx950_code(fem):
    """The offload bundle of the library built here names its target."""
    data = open(fem.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_a_missing_library_fails_loudly(fem, monkeypatch):
    monkeypatch.setattr(fem, "_lib", None)
    monkeypatch.setattr(fem, "LIB_PATH", os.path.join(ROOT, "no_such_library.so"))
    with pytest.raises(fem.FemError):
        fem.lib()
