"""Pack the reference's viscoelastic attenuation benchmark (EXAMPLES/benchmarks/attenuation/viscoelastic)
into one fixture:

* its closed-form solution (REF_ANALYTICAL/U?_time_analytical_solution_viscoelastic.dat and the
  elastic one): displacement 500 m x (1, 1, 1) away from a vertical unit force with an 18 Hz Ricker
  STF in a full space with rho 2000, vp 3297.849, vs 2222.536 at 18 Hz, Q_kappa 10, Q_mu 20 (the
  first 0.6 s, time axis t = (it-1) dt - 1.2/f0 as the solver's);
* the reference's own run of it (REF_SEIS/CE.1.FX?.semd, 1300 steps of 0.5 ms, 80^3 elements);
* the SLS constants that closed form was computed with, as listed in
  analytical_solution_viscoelastic_3D_Carcione_and_Komatitsch_correct_with_1_over_L.f90
  (tau_sigma, tau_epsilon for Q_kappa and Q_mu, factor_scale_kappa/mu);
* DATA/FORCESOLUTION and DATA/STATIONS text.

    python tests/golden/make_attenuation_benchmark_fixture.py <path to the benchmark directory>
"""
import os
import re
import sys

import numpy as np


def fortran_values(text, name):
    m = re.search(name + r"\s*=\s*\(/(.*?)/\)", text, re.S)
    return np.array([float(v.replace("d", "e")) for v in m.group(1).replace("&", " ").split(",")])


def fortran_scalar(text, name):
    m = re.search(r"parameter\s*::\s*" + name + r"\s*=\s*([0-9.d+-]+)", text)
    return float(m.group(1).replace("d", "e"))


if __name__ == "__main__":
    ex = sys.argv[1]
    out = {}
    for c in "xyz":
        for kind in ("viscoelastic", "elastic"):
            a = np.loadtxt(os.path.join(ex, "REF_ANALYTICAL", f"U{c}_time_analytical_solution_{kind}.dat"))
            a = a[a[:, 0] <= 0.6]
            out[f"analytical_{kind}_{c}"] = a.astype(np.float64)
    seis = np.stack([np.loadtxt(os.path.join(ex, "REF_SEIS", f"CE.1.FX{c}.semd")) for c in "XYZ"], 1)
    out["ref_seis"] = seis[:, :, 1].astype(np.float32)
    out["t"] = seis[:, 0, 0].astype(np.float32)
    src = open(os.path.join(ex, "analytical_solution_viscoelastic_3D_Carcione_and_Komatitsch_correct_with_1_over_L.f90")).read()
    for nm in ("tau_epsilon_kappa", "tau_sigma_kappa", "tau_epsilon_mu", "tau_sigma_mu"):
        out[nm] = fortran_values(src, nm)
    out["factor_scale_mu"] = np.array(fortran_scalar(src, "factor_scale_mu"))
    out["factor_scale_kappa"] = np.array(fortran_scalar(src, "factor_scale_kappa"))
    out["forcesolution"] = np.array(open(os.path.join(ex, "DATA", "FORCESOLUTION")).read())
    out["stations"] = np.array(open(os.path.join(ex, "DATA", "STATIONS")).read())
    out["DT"] = np.array(0.0005)
    out["NSTEP"] = np.array(1300)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "attenuation_benchmark.npz"), **out)
    print({k: (v.shape if v.ndim else v) for k, v in out.items() if k.startswith(("tau", "factor"))})
