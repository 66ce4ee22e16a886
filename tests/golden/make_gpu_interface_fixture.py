"""Record the reference's GPU interface as data: every Fortran-callable GPU routine with its C parameter
types (src/gpu/specfem3D_gpu_cuda_method_stubs.c, the list the reference links when built without a
GPU) and every call of those routines in the Fortran solver sources (src/specfem3D, src/shared),
with its argument count and file:line. A Fortran host calls them through implicit interfaces, so a
drop-in library must export each name with the same number of (by-reference) arguments.

    python tests/golden/make_gpu_interface_fixture.py <path to the SPECFEM3D source tree>

writes tests/golden/reference_gpu_interface.json.
"""
import json
import os
import re
import sys


def stub_signatures(path):
    text = open(path).read()
    text = re.sub(r"//[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"FC_FUNC_\(\s*(\w+)\s*,\s*\w+\s*\)\s*\(([^)]*)\)", text):
        name, params = m.group(1), m.group(2)
        types = []
        for p in params.split(","):
            p = " ".join(p.replace("const", " ").split())
            if not p or p == "void":
                continue
            ptr = p.count("*")
            base = p.replace("*", " ").split()[0]
            types.append({"base": base, "pointer": ptr})
        out[name.lower()] = types
    return out


def strip_comment(line):
    # Fortran free-form comment: '!' outside a string
    q = None
    for i, ch in enumerate(line):
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "!":
            return line[:i]
    return line


def fortran_calls(root, names):
    calls = {}
    pat = re.compile(r"\bcall\s+(\w+)\s*\(", re.I)
    for sub in ("src/specfem3D", "src/shared"):
        d = os.path.join(root, sub)
        for fn in sorted(os.listdir(d)):
            if not fn.lower().endswith((".f90", ".f")):
                continue
            lines = open(os.path.join(d, fn), errors="replace").read().split("\n")
            # join continuation lines, remembering the first line number
            stmts, cur, start = [], "", 0
            for ln, raw in enumerate(lines, 1):
                s = strip_comment(raw).rstrip()
                if not cur:
                    start = ln
                st = s.lstrip()
                if st.startswith("&"):
                    st = st[1:]
                if st.endswith("&"):
                    cur += st[:-1] + " "
                    continue
                cur += st
                stmts.append((start, cur))
                cur = ""
            for ln, st in stmts:
                for m in pat.finditer(st):
                    name = m.group(1).lower()
                    if name not in names:
                        continue
                    depth, i, args, tok = 1, m.end(), 0, ""
                    while i < len(st) and depth:
                        ch = st[i]
                        if ch == "(":
                            depth += 1
                        elif ch == ")":
                            depth -= 1
                            if depth == 0:
                                break
                        elif ch == "," and depth == 1:
                            args += 1
                            tok = ""
                            i += 1
                            continue
                        tok += ch
                        i += 1
                    nargs = args + (1 if (args or tok.strip()) else 0)
                    calls.setdefault(name, []).append({"nargs": nargs, "at": f"{sub}/{fn}:{ln}"})
    return calls


if __name__ == "__main__":
    root = sys.argv[1]
    stubs = stub_signatures(os.path.join(root, "src/gpu/specfem3D_gpu_cuda_method_stubs.c"))
    calls = fortran_calls(root, set(stubs))
    here = os.path.dirname(os.path.abspath(__file__))
    out = {"source": "src/gpu/specfem3D_gpu_cuda_method_stubs.c; calls in src/specfem3D, src/shared",
           "routines": stubs, "fortran_calls": calls}
    with open(os.path.join(here, "reference_gpu_interface.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print(len(stubs), "routines,", sum(len(v) for v in calls.values()), "call sites of", len(calls), "routines")
