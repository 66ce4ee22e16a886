"""Pack the reference example ``homogeneous_halfspace_HEX8_elastic_absorbing_Stacey_5sides``
(36 x 36 x 16 irregular HEX8 elements, Stacey on 5 sides, free top) into one fixture.

Everything here is data the reference ships with its own regression tests (EXAMPLES/applications/
homogeneous_halfspace_HEX8_elastic_absorbing_Stacey_5sides/{MESH-default,DATA,REF_SEIS}):
the CUBIT-format mesh arrays, the CMTSOLUTION and STATIONS values and the reference seismograms
(REF_SEIS/DB.X??.BX?.semd, 5000 samples, DT = 0.05 s).

    python tests/golden/make_stacey5_fixture.py <path to the example directory>
"""
import os
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
    # nummaterial_velocity_file: domain_id material_id rho vp vs Q_kappa Q_mu anisotropy
    rows = []
    for line in open(os.path.join(m, "nummaterial_velocity_file")):
        p = line.split("!")[0].split()
        if len(p) >= 7 and p[0].lstrip("-").isdigit():
            rows.append([float(v) for v in p[:7]])
    out["material_props"] = np.array(rows)
    out["cmtsolution"] = np.array(open(os.path.join(ex, "DATA", "CMTSOLUTION")).read())
    out["stations"] = np.array(open(os.path.join(ex, "DATA", "STATIONS")).read())
    st = [l.split() for l in open(os.path.join(ex, "DATA", "STATIONS")) if l.split()]
    seis = np.zeros((5000, len(st), 3), np.float32)
    for r, s in enumerate(st):
        for c, ch in enumerate("XYZ"):
            a = np.loadtxt(os.path.join(ex, "REF_SEIS", f"{s[1]}.{s[0]}.BX{ch}.semd"))
            seis[:, r, c] = a[:, 1]
            t = a[:, 0]
    out["ref_seis"] = seis
    out["t"] = t.astype(np.float32)
    out["DT"] = np.array(0.05)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "stacey5_example.npz"), **out)
    print("wrote", os.path.join(here, "stacey5_example.npz"))
