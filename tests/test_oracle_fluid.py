"""Sources and receivers in fluid elements: host set-up and the oracle, pinned by closed forms.

* STFs: the pressure-trick STFs are the second time derivatives of the Gaussian and Ricker STFs
  (central differences), and a Gaussian source in a fluid starts the run at t0 = -3 (tshift - hdur);
* source array: an explosive CMT (Mxx = Myy = Mzz = M) in a fluid has the scalar factor M, spread by
  Lagrange weights that sum to one;
* receivers: on a linear potential the pressure interpolates to -chi(x_r) and the vector record is
  (1/rho) grad chi, exactly, in regular and warped (irregular) elements;
* time stepping: in a closed fluid box the mass-weighted sum of chi'' equals the injected source,
  -(factor / kappa) stf(t_n), at every step (the stiffness rows sum to zero); the same for an adjoint
  source at a fluid receiver (+ adj_n). A sign or 1/kappa mistake planted in the oracle's source term
  breaks it.
"""
import math

import numpy as np
import pytest

from conftest import smooth_warp


def fluid_box(n=3, warp=None, dt=0.002, stacey=False, free=False):
    from specfem3d_amd.mesh.box import build_box, homogeneous
    L = 500.0 * n
    ext = (0.0, L, 0.0, L, -L, 0.0)
    db = build_box(n, n, n, ext, material=homogeneous(1020.0, 1500.0, 0.0), acoustic_top_layers=n,
                   acoustic_material=homogeneous(1020.0, 1500.0, 0.0),
                   stacey_faces=(1, 2, 3, 4, 5) if stacey else (), free_surface=free,
                   warp=warp(L) if warp else None)
    db.finalize(dt, stacey=stacey)
    assert db.nspec_acoustic == db.nspec
    return db, L


def test_pressure_trick_stfs_are_second_derivatives():
    from specfem3d_amd import sources as S
    t = np.linspace(-3.0, 3.0, 601)
    h = 1e-4
    for f, d2, p in ((S.stf_gaussian, S.stf_d2gaussian, 0.7), (S.stf_ricker, S.stf_d2ricker, 1.3)):
        num = (f(t + h, p) - 2.0 * f(t, p) + f(t - h, p)) / h ** 2
        ana = d2(t, p)
        assert np.abs(num - ana).max() < 1e-5 * np.abs(ana).max()


def test_fluid_source_array_and_t0():
    from specfem3d_amd import sources as S
    db, L = fluid_box(n=3)
    M = 2.5e15
    cmt = S.CMTSource(x=0.41 * L, y=0.52 * L, depth_km=0.37 * L / 1000.0, tshift=0.0, hdur=0.05,
                      Mxx=M, Myy=M, Mzz=M, Mxy=0.0, Mxz=0.0, Myz=0.0)
    assert math.isclose(S.cmt_acoustic_factor(cmt), M, rel_tol=1e-14)
    src = S.setup_cmt_sources(db, [cmt], db.dt, pressure_trick=True)
    assert bool(src.acoustic[0]) and src.kind_acoustic == "d2gaussian"
    assert math.isclose(src.t0, 3.0 * 0.05)                 # -3 (tshift - hdur) beats -2 (tshift - hdur)
    a = src.sourcearrays[0].astype(np.float64)
    assert np.allclose(a[:, 0], a[:, 1]) and np.allclose(a[:, 0], a[:, 2])
    assert math.isclose(a[:, 0].sum(), M, rel_tol=1e-6)
    tab = src.stf_table(40)
    t = (np.arange(40) * db.dt) - src.t0
    assert np.allclose(tab[:, 0], S.stf_d2gaussian(t, 0.05 / S.SOURCE_DECAY_MIMIC_TRIANGLE))


@pytest.mark.parametrize("warp", [False, True])
def test_fluid_receivers_exact_on_linear_potential(oracle_mod, warp):
    from specfem3d_amd import sources as S
    db, L = fluid_box(n=3, warp=smooth_warp if warp else None)
    assert (db.nspec_irregular > 0) == warp
    rng = np.random.default_rng(7)
    st = [S.Station(f"R{i}", "T", float(rng.uniform(0.1, 0.9) * L), float(rng.uniform(0.1, 0.9) * L), 0.0,
                    float(rng.uniform(0.1, 0.9) * L)) for i in range(8)]
    rec = S.setup_receivers(db, st)
    c = np.array([0.3, -1.1, 0.7])
    chi = (5.0 + c[0] * db.xstore.astype(np.float64) + c[1] * db.ystore + c[2] * db.zstore)
    o = oracle_mod.OracleSolver(db, None, rec, "f64")
    o.pot[:] = chi
    o.potdd[:] = chi
    xr = np.array([[Lc.x, Lc.y, Lc.z] for Lc in rec.located])
    expect_p = -(5.0 + xr @ c)
    # the interpolants are CUSTOM_REAL (float32) and the coordinates float32: 1e-6 / 1e-5 tolerances
    for kind in ("pressure", "pressure_trick"):
        p = o.seismogram(kind)
        assert np.allclose(p[:, 0], expect_p, rtol=1e-6, atol=1e-6 * np.abs(expect_p).max())
        assert np.all(p[:, 1:] == 0)
    d = o.seismogram("displ")
    assert np.allclose(d, np.tile(c / 1020.0, (rec.n, 1)), rtol=1e-5, atol=1e-5 * np.abs(c).max() / 1020.0)


def _mass_sum(o):
    M = 1.0 / o.rmass_ac.astype(np.float64)
    return float((M * o.potdd.astype(np.float64)).sum())


def test_fluid_source_injects_its_stf(oracle_mod):
    """Closed fluid box (no Stacey, no free surface): sum_g M_g chi''_g = -(factor/kappa) stf(t_n)."""
    from specfem3d_amd import sources as S
    db, L = fluid_box(n=3, warp=smooth_warp)
    M = 1.0e12
    cmt = S.CMTSource(x=0.43 * L, y=0.55 * L, depth_km=0.48 * L / 1000.0, tshift=0.0, hdur=0.04,
                      Mxx=M, Myy=M, Mzz=M, Mxy=0.0, Mxz=0.0, Myz=0.0)
    src = S.setup_cmt_sources(db, [cmt], db.dt, pressure_trick=False)
    o = oracle_mod.OracleSolver(db, src, None, "f64")
    o.stf_table = src.stf_table(60)
    kappa = 1020.0 * 1500.0 ** 2
    err = []
    for it in range(1, 61):
        o.step()
        want = -(src.sourcearrays[0, :, 0].astype(np.float64).sum() / kappa) * o.stf_table[it - 1, 0]
        err.append(abs(_mass_sum(o) - want) / (abs(want) + 1e-300))
    # float32 tables (hprime rows, inverse mass) leave ~1e-7; a sign or 1/kappa slip is O(1) or worse
    assert max(err) < 1e-6, max(err)


def test_fluid_adjoint_source_injects_its_trace(oracle_mod):
    from specfem3d_amd import sources as S
    db, L = fluid_box(n=3, warp=smooth_warp)
    st = [S.Station("A", "T", 0.46 * L, 0.38 * L, 0.0, 0.51 * L), S.Station("B", "T", 0.7 * L, 0.3 * L, 0.0, 0.2 * L)]
    rec = S.setup_receivers(db, st)
    o = oracle_mod.OracleSolver(db, None, rec, "f64")
    rng = np.random.default_rng(3)
    for it in range(20):
        adj = rng.normal(size=(2, 3))
        o.step(adjoint_source=adj)
        # float interpolants' products: sum to one within float rounding
        assert abs(_mass_sum(o) - adj[:, 0].sum()) < 1e-6 * np.abs(adj[:, 0]).sum()
