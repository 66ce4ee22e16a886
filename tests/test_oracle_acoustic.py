"""Pins of the oracle's acoustic path (compute_forces_acoustic.F90) and of the fluid-solid
coupling signs (compute_coupling_*.f90) against mathematics.

* the acoustic stiffness annihilates constant potentials and is symmetric;
* a closed fluid box conserves the discrete energy;
* in a closed coupled fluid-over-solid box the semi-discrete energy
  E = 1/2 u'M u' + 1/2 u K u + 1/2 chi''M_f chi'' + 1/2 chi' K_f chi' is conserved only with the
  reference's coupling signs (chi'' += w u.n, a += -w n chi''); flipping either sign makes it drift.
"""
import numpy as np
import pytest

from conftest import rel_l2
from specfem3d_amd.mesh.box import build_box, homogeneous


def fluid_box(n=3, warp=None):
    L = 500.0 * n
    db = build_box(n, n, n, (0, L, 0, L, -L, 0), acoustic_top_layers=n, acoustic_material=homogeneous(1020, 1500, 0),
                   stacey_faces=(), free_surface=False, warp=warp)
    db.finalize(0.004, stacey=False)
    return db


def ac_force(oracle_mod, db, pot):
    o = oracle_mod.OracleSolver(db, None, None, "f64")
    o.pot[:] = pot
    o.potdd[:] = 0
    o.forces_acoustic()
    return o.potdd.copy()


def test_acoustic_stiffness_constant_and_symmetry(oracle_mod):
    from conftest import smooth_warp
    db = fluid_box(3, warp=smooth_warp(1500.0))
    n = db.nglob
    rng = np.random.default_rng(0)
    scale = np.abs(ac_force(oracle_mod, db, rng.normal(size=n))).max()
    assert np.abs(ac_force(oracle_mod, db, np.full(n, 3.0))).max() < 1e-6 * scale
    u, v = rng.normal(size=n), rng.normal(size=n)
    a, b = v @ ac_force(oracle_mod, db, u), u @ ac_force(oracle_mod, db, v)
    assert abs(a - b) < 1e-6 * abs(a)


def test_closed_fluid_box_conserves_energy(oracle_mod):
    db = fluid_box(3)
    o = oracle_mod.OracleSolver(db, None, None, "f64")
    x, y, z = (db.xstore.astype(np.float64), db.ystore.astype(np.float64), db.zstore.astype(np.float64))
    o.pot[:] = np.cos(x / 300.0) * np.sin(y / 400.0 + z / 500.0)
    M = 1.0 / db.rmass_acoustic.astype(np.float64)
    E, prev = [], o.pot.copy()
    for _ in range(300):
        o.step()
        vh = o.potd - o.deltatover2 * o.potdd
        E.append(0.5 * vh @ (M * vh) - 0.5 * prev @ (M * o.potdd))
        prev = o.pot.copy()
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-6 * abs(E[0])


def coupled_energy(oracle_mod, flip_ac=False, flip_el=False, nstep=600):
    L = 1500.0
    db = build_box(3, 3, 3, (0, L, 0, L, -L, 0), material=homogeneous(2500, 3500, 2000), acoustic_top_layers=1,
                   acoustic_material=homogeneous(1000, 1500, 0), stacey_faces=(), free_surface=False)
    db.finalize(0.002, stacey=False)
    o = oracle_mod.OracleSolver(db, None, None, "f64")
    # plant a sign mistake in one coupling direction of the oracle (test of the test)
    o.plant_flip_ac_el = int(flip_ac)
    o.plant_flip_el_ac = int(flip_el)
    x, y, z = (db.xstore.astype(np.float64), db.ystore.astype(np.float64), db.zstore.astype(np.float64))
    r2 = (x - 700) ** 2 + (y - 800) ** 2 + (z + 800) ** 2
    o.displ[:] = (np.exp(-r2 / 250.0 ** 2)[:, None] * np.array([1.0, 0.5, -0.7])).ravel() * 1e-3
    Mel = 1.0 / np.stack([db.rmassx, db.rmassy, db.rmassz], 1).astype(np.float64).ravel()
    Mac = 1.0 / db.rmass_acoustic.astype(np.float64)
    elmask = np.zeros(db.nglob, bool)
    elmask[np.unique(db.ibool[db.ispec_is_elastic])] = True
    acmask = np.zeros(db.nglob, bool)
    acmask[np.unique(db.ibool[db.ispec_is_acoustic])] = True
    Mel = Mel * np.repeat(elmask, 3)
    Mac = Mac * acmask
    K = oracle_mod.OracleSolver(db, None, None, "f64")  # stiffness actions without coupling

    def Kel(u):
        K.displ[:] = u
        K.accel[:] = 0
        K.forces_elastic()
        return -K.accel

    def Kac(c):
        K.pot[:] = c
        K.potdd[:] = 0
        K.forces_acoustic()
        return -K.potdd

    E = []
    for _ in range(nstep):
        o.step()
        ke = 0.5 * o.veloc @ (Mel * o.veloc)
        pe = 0.5 * o.displ @ Kel(o.displ)
        kf = 0.5 * o.potdd @ (Mac * o.potdd) + 0.5 * o.potd @ Kac(o.potd)
        E.append(ke + pe + kf)
    return np.array(E)


def test_coupled_energy_and_sign_sensitivity(oracle_mod):
    E = coupled_energy(oracle_mod)
    drift = np.abs(E - E[0]).max() / E[0]
    assert drift < 0.05, drift
    for kw in (dict(flip_ac=True), dict(flip_el=True)):
        Ebad = coupled_energy(oracle_mod, **kw)
        bad = np.abs(Ebad - Ebad[0]).max() / Ebad[0]
        assert bad > 10 * drift, (kw, drift, bad)
