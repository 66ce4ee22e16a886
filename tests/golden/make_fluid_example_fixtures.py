"""Pack the reference's two fluid examples into fixtures:

* ``homogeneous_acoustic`` (36 x 36 x 16 acoustic HEX8 elements, Stacey on 5 sides, explosive CMT
  source in the fluid, pressure seismograms with USE_TRICK_FOR_BETTER_PRESSURE, 1000 steps of
  DT = 0.1136 s): REF_SEIS/DB.X??.MXP.semp;
* ``waterlayered_halfspace`` (40,500 elements: a 3 km water layer over an elastic half-space,
  Stacey on 5 sides, explosive CMT source in the water, 4500 steps of DT = 0.005 s):
  REF_SEIS/DB.*.HXZ.sem{d,v,a} (vertical components of displacement, velocity, acceleration).

Everything packed is data the reference ships for its own regression tests
(EXAMPLES/applications/<example>/{MESH-default,DATA,REF_SEIS}): the CUBIT-format mesh arrays,
CMTSOLUTION and STATIONS text, and the reference seismograms (missing traces are NaN).

    python tests/golden/make_fluid_example_fixtures.py <path to EXAMPLES/applications>
"""
import os
import sys

import numpy as np


def table(path, dtype=float):
    with open(path) as f:
        n = int(f.readline().split()[0])
        return np.loadtxt(f, dtype=dtype, max_rows=n, ndmin=2)


def mesh_arrays(ex):
    m = os.path.join(ex, "MESH-default")
    out = {"nodes": table(os.path.join(m, "nodes_coords_file")),
           "elements": table(os.path.join(m, "mesh_file"), np.int64).astype(np.int32),
           "materials": np.loadtxt(os.path.join(m, "materials_file"), dtype=np.int32, ndmin=2)}
    for nm in ("xmin", "xmax", "ymin", "ymax", "bottom"):
        out["abs_" + nm] = table(os.path.join(m, f"absorbing_surface_file_{nm}"), np.int64).astype(np.int32)
    out["free_zmax"] = table(os.path.join(m, "free_or_absorbing_surface_file_zmax"), np.int64).astype(np.int32)
    rows = []
    for line in open(os.path.join(m, "nummaterial_velocity_file")):
        p = line.split("!")[0].split()
        if len(p) >= 7 and p[0].lstrip("-").isdigit():
            rows.append([float(v) for v in p[:7]])
    out["material_props"] = np.array(rows)
    out["cmtsolution"] = np.array(open(os.path.join(ex, "DATA", "CMTSOLUTION")).read())
    out["stations"] = np.array(open(os.path.join(ex, "DATA", "STATIONS")).read())
    return out


def seis(ex, st, nstep, names):
    """(nstep, nsta, len(names)) float32 traces REF_SEIS/NET.STA.<name>, NaN where absent; time column."""
    out = np.full((nstep, len(st), len(names)), np.nan, np.float32)
    t = None
    for r, s in enumerate(st):
        for c, nm in enumerate(names):
            p = os.path.join(ex, "REF_SEIS", f"{s[1]}.{s[0]}.{nm}")
            if os.path.exists(p):
                a = np.loadtxt(p)
                out[:, r, c] = a[:nstep, 1]
                t = a[:nstep, 0]
    return out, t


if __name__ == "__main__":
    apps = sys.argv[1]
    here = os.path.dirname(os.path.abspath(__file__))

    ex = os.path.join(apps, "homogeneous_acoustic")
    out = mesh_arrays(ex)
    st = [l.split() for l in open(os.path.join(ex, "DATA", "STATIONS")) if l.split()]
    out["ref_seis"], t = seis(ex, st, 1000, ["MXP.semp"])
    out["t"] = t.astype(np.float32)
    out["DT"] = np.array(0.1136)
    out["NSTEP"] = np.array(1000)
    np.savez_compressed(os.path.join(here, "homogeneous_acoustic_example.npz"), **out)

    ex = os.path.join(apps, "waterlayered_halfspace")
    out = mesh_arrays(ex)
    st = [l.split() for l in open(os.path.join(ex, "DATA", "STATIONS")) if l.split()]
    out["ref_seis"], t = seis(ex, st, 4500, ["HXZ.semd", "HXZ.semv", "HXZ.sema"])
    out["t"] = t.astype(np.float32)
    out["DT"] = np.array(0.005)
    out["NSTEP"] = np.array(4500)
    np.savez_compressed(os.path.join(here, "waterlayered_halfspace_example.npz"), **out)
    print("wrote the two fixtures into", here)
