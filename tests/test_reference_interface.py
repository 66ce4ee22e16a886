"""The library's C-ABI against the reference's GPU interface, recorded as data by
tests/golden/make_gpu_interface_fixture.py (no reference file is read here):

* every Fortran-callable GPU routine of the reference (src/gpu/specfem3D_gpu_cuda_method_stubs.c) is
  declared in include/specfem3d_amd.h with the trailing-underscore name, the same number of
  parameters and compatible types (pointer depth and base type; `field` is `realw`, const ignored);
* every call of those routines in the reference's Fortran solver (implicit interfaces: the argument
  count is the whole contract) passes as many arguments as the library's declaration takes;
* the built library exports them all.

Routines outside the single-partition hot path are exported too and stop the run with a message.
"""
import json
import os
import re

from conftest import GOLDEN, ROOT

def header_decls():
    src = open(os.path.join(ROOT, "include", "specfem3d_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:void|int|long|double|const char\s*\*)\s+(\w+)\s*\(([^)]*)\)\s*;", src):
        types = []
        for p in m.group(2).split(","):
            p = " ".join(p.replace("const", " ").split())
            if not p or p == "void":
                continue
            ptr = p.count("*")
            types.append({"base": p.replace("*", " ").split()[0], "pointer": ptr})
        out[m.group(1)] = types
    return out


def fixture():
    return json.load(open(os.path.join(GOLDEN, "reference_gpu_interface.json")))


def norm(t):
    b = {"field": "realw"}.get(t["base"], t["base"])
    return b, t["pointer"]


def test_every_reference_routine_declared_with_same_parameters():
    ref = fixture()["routines"]
    ours = header_decls()
    missing = [n for n in ref if n + "_" not in ours]
    assert not missing, missing
    bad = []
    for n, params in ref.items():
        mine = ours[n + "_"]
        if len(mine) != len(params):
            bad.append((n, "count", len(params), len(mine)))
            continue
        for k, (a, b) in enumerate(zip(params, mine)):
            if norm(a) != norm(b):
                bad.append((n, k, norm(a), norm(b)))
    assert not bad, bad


def test_fortran_call_sites_match_argument_counts():
    fx = fixture()
    ours = header_decls()
    assert sum(len(v) for v in fx["fortran_calls"].values()) > 150
    bad = [(n, c["at"], c["nargs"], len(ours[n + "_"])) for n, cs in fx["fortran_calls"].items() for c in cs
           if len(ours[n + "_"]) != c["nargs"]]
    assert not bad, bad


def test_library_exports_reference_routines():
    from specfem3d_amd import abi
    lib = abi.load()
    missing = [n for n in fixture()["routines"] if not hasattr(lib, n + "_")]
    assert not missing, missing
