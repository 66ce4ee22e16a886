"""Pack the reference example ``CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides``
(36 x 36 x 16 HEX8 elements, C-PML layers three elements thick on the four sides and the bottom, free
top, homogeneous elastic rho 2300, vp 2800, vs 1500) into one fixture.

Everything here is data the reference ships with this example (MESH-default, DATA, REF_SEIS):

* the CUBIT-format mesh arrays, including ``absorbing_cpml_file`` (element, region code);
* DATA/FORCESOLUTION and DATA/STATIONS text; DT = 0.11 s, NSTEP = 2400, f0_FOR_PML = 0.05555 Hz
  (DATA/Par_file);
* the reference seismograms REF_SEIS/DB.X??.MX?.semd (displacement, 2400 samples);
* values its generate_databases and solver logs print: the C-PML widths and interface origins
  (output_generate_databases.txt) and the maximum displacement norm every 100 steps
  (output_solver.txt, "Max norm displacement vector U").

    python tests/golden/make_cpml5_fixture.py <path to the example directory>
"""
import os
import re
import sys

import numpy as np


def table(path, dtype=float):
    with open(path) as f:
        n = int(f.readline().split()[0])
        return np.loadtxt(f, dtype=dtype, max_rows=n, ndmin=2)


if __name__ == "__main__":
    ex = sys.argv[1]
    m = os.path.join(ex, "MESH-default")
    out = {}
    out["nodes"] = table(os.path.join(m, "nodes_coords_file"))
    out["elements"] = table(os.path.join(m, "mesh_file"), np.int64).astype(np.int32)
    out["materials"] = np.loadtxt(os.path.join(m, "materials_file"), dtype=np.int32, ndmin=2)
    for nm in ("xmin", "xmax", "ymin", "ymax", "bottom"):
        out["abs_" + nm] = table(os.path.join(m, f"absorbing_surface_file_{nm}"), np.int64).astype(np.int32)
    out["free_zmax"] = table(os.path.join(m, "free_or_absorbing_surface_file_zmax"), np.int64).astype(np.int32)
    out["cpml"] = table(os.path.join(m, "absorbing_cpml_file"), np.int64).astype(np.int32)
    rows = []
    for line in open(os.path.join(m, "nummaterial_velocity_file")):
        p = line.split("!")[0].split()
        if len(p) >= 7 and p[0].lstrip("-").isdigit():
            rows.append([float(v) for v in p[:7]])
    out["material_props"] = np.array(rows)
    out["forcesolution"] = np.array(open(os.path.join(ex, "DATA", "FORCESOLUTION")).read())
    out["stations"] = np.array(open(os.path.join(ex, "DATA", "STATIONS")).read())
    par = open(os.path.join(ex, "DATA", "Par_file")).read()
    val = lambda k: re.search(r"^" + k + r"\s*=\s*(\S+)", par, re.M).group(1)
    out["DT"] = np.array(float(val("DT")))
    out["NSTEP"] = np.array(int(val("NSTEP")))
    out["f0_FOR_PML"] = np.array(float(val("f0_FOR_PML")))
    st = [l.split() for l in open(os.path.join(ex, "DATA", "STATIONS")) if l.split()]
    nstep = int(out["NSTEP"])
    seis = np.zeros((nstep, len(st), 3), np.float32)
    for r, s in enumerate(st):
        for c, ch in enumerate("XYZ"):
            a = np.loadtxt(os.path.join(ex, "REF_SEIS", f"{s[1]}.{s[0]}.MX{ch}.semd"))
            seis[:, r, c] = a[:, 1]
            t = a[:, 0]
    out["ref_seis"] = seis
    out["t"] = t.astype(np.float32)
    log = open(os.path.join(ex, "REF_SEIS", "output_generate_databases.txt")).read()
    w = [float(re.search(r"CPML_width_" + c + r":\s*(\S+)", log).group(1)) for c in "xyz"]
    out["log_cpml_width"] = np.array(w)
    o = re.search(r"Origins of right/left X-surface C-PML\s+(\S+)\s+(\S+)", log)
    oy = re.search(r"Origins of front/back Y-surface C-PML\s+(\S+)\s+(\S+)", log)
    oz = re.search(r"Origin of bottom Z-surface C-PML\s+(\S+)", log)
    out["log_cpml_origins"] = np.array([float(o.group(1)), float(o.group(2)), float(oy.group(1)),
                                        float(oy.group(2)), float(oz.group(1))])
    out["log_nspec_cpml"] = np.array(int(re.search(r"total number of C-PML elements in the global mesh:\s*(\d+)",
                                                   log).group(1)))
    sol = open(os.path.join(ex, "REF_SEIS", "output_solver.txt")).read()
    steps = [int(v) for v in re.findall(r"Time step #\s+(\d+)", sol)]
    norms = [float(v) for v in re.findall(r"Max norm displacement vector U in all slices \(m\) =\s*(\S+)", sol)]
    out["log_norm_steps"] = np.array(steps[:len(norms)], np.int32)
    out["log_norms"] = np.array(norms)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "cpml5_example.npz"), **out)
    print("wrote", os.path.join(here, "cpml5_example.npz"), {k: np.shape(v) for k, v in out.items()})
