"""The viscoelastic (3-SLS) path against the reference benchmark's closed-form solution.

EXAMPLES/benchmarks/attenuation/viscoelastic: a vertical unit force with an 18 Hz Ricker STF in a
homogeneous viscoelastic medium with Q_kappa = 10 and Q_mu = 20, receiver 500 m x (1, 1, 1) away;
REF_ANALYTICAL holds the closed-form displacement (Carcione's solution with the SLS constants the
benchmark lists) and REF_SEIS the reference's own run. The GPU run (FMA build, atomic assembly)
goes through the whole attenuated path: host SLS set-up, scaled moduli, memory-variable update and
stress correction in the stiffness kernel, point force, receiver.

Tolerances: against the closed form 6e-3 aggregate rel-L2 over the three components (the
reference's own run is 5.08e-3 away, and so is this one; the purely elastic closed form is 0.69
away over the same window, and a 1 % change of 1/Q moves the traces by about 1e-2); against the
reference's run 1e-5 (measured on MI355X: 1.0e-6).
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_attenuation_benchmark_closed_form():
    from scipy.interpolate import CubicSpline
    from specfem3d_amd.examples import attenuation_benchmark_example
    from specfem3d_amd.solver import GPUSolver
    db, src, rec, att, fx = attenuation_benchmark_example()
    nstep = int(fx["NSTEP"])
    print(f"[gpu] attenuation benchmark: nspec={db.nspec} nglob={db.nglob} nstep={nstep}", flush=True)
    g = GPUSolver(db, src, rec, nstep=nstep, attenuation=att, exact=False, assembly=0)
    out = g.run()[:, 0, :].astype(np.float64)
    g.close()
    t = np.arange(nstep) * db.dt - src.t0
    assert np.allclose(t, fx["t"], atol=1e-6)
    ana, ela = [], []
    a0 = fx["analytical_viscoelastic_x"][0, 0]
    m = t >= a0
    for c in "xyz":
        a = fx[f"analytical_viscoelastic_{c}"]
        e = fx[f"analytical_elastic_{c}"]
        ana.append(CubicSpline(a[:, 0], a[:, 1])(t[m]))
        ela.append(CubicSpline(e[:, 0], e[:, 1])(t[m]))
    ana, ela = np.stack(ana, 1), np.stack(ela, 1)
    e_ana = rel_l2(out[m], ana)
    e_ref = rel_l2(out, fx["ref_seis"])
    e_ela = rel_l2(out[m], ela)
    print(f"[gpu] attenuation benchmark: rel-L2 vs closed form {e_ana:.3e} (reference run "
          f"{rel_l2(fx['ref_seis'][m], ana):.3e}), vs reference run {e_ref:.3e}, vs elastic closed form {e_ela:.3f}",
          flush=True)
    assert e_ana < 6e-3
    assert e_ref < 1e-5
    assert e_ela > 0.5
