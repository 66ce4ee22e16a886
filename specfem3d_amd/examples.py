"""Reference examples packed as fixtures (inputs + reference seismograms)."""
from __future__ import annotations

import io
import os

import numpy as np

from . import sources as S
from .mesh.external import build_from_arrays

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def stacey5_example(path: str | None = None):
    """(db, sources, receivers, ref_seis, DT) of the reference example
    homogeneous_halfspace_HEX8_elastic_absorbing_Stacey_5sides (CMT with error-function STF)."""
    z = np.load(path or os.path.join(GOLDEN, "stacey5_example.npz"), allow_pickle=False)
    dt = float(z["DT"])
    db = build_from_arrays({k: z[k] for k in z.files}, dt)
    cmts = S.read_cmtsolution(io.StringIO(str(z["cmtsolution"])), dt)
    # this example's REF_SEIS come from an NPROC=4 run whose partition put the corner source
    # in the element below it
    src = S.setup_cmt_sources(db, cmts, dt, ricker=False, tie_break="lower")
    st = S.read_stations(io.StringIO(str(z["stations"])))
    rec = S.setup_receivers(db, st)
    return db, src, rec, z["ref_seis"], dt


def _fluid_example(name: str, path: str | None, pressure_trick: bool):
    z = np.load(path or os.path.join(GOLDEN, f"{name}_example.npz"), allow_pickle=False)
    dt = float(z["DT"])
    db = build_from_arrays({k: z[k] for k in z.files}, dt)
    cmts = S.read_cmtsolution(io.StringIO(str(z["cmtsolution"])), dt)
    src = S.setup_cmt_sources(db, cmts, dt, ricker=False, pressure_trick=pressure_trick)
    st = S.read_stations(io.StringIO(str(z["stations"])))
    rec = S.setup_receivers(db, st)
    return db, src, rec, z["ref_seis"], dt, int(z["NSTEP"]), st


def homogeneous_acoustic_example(path: str | None = None):
    """(db, sources, receivers, ref_pressure (nstep, nrec, 1), DT, NSTEP, stations) of the reference example
    homogeneous_acoustic: a fluid-only CUBIT mesh with Stacey on 5 sides, an explosive CMT source in
    the fluid with the Gaussian STF's second derivative (USE_TRICK_FOR_BETTER_PRESSURE) and pressure
    receivers 50 m below the free surface."""
    return _fluid_example("homogeneous_acoustic", path, pressure_trick=True)


def waterlayered_example(path: str | None = None):
    """(db, sources, receivers, ref (nstep, nrec, 3) vertical displacement/velocity/acceleration with NaN
    where the example has no trace, DT, NSTEP, stations) of the reference example waterlayered_halfspace:
    a 3 km water layer over an elastic half-space, explosive CMT source in the water (Gaussian STF),
    receivers in both domains."""
    return _fluid_example("waterlayered_halfspace", path, pressure_trick=False)


def attenuation_benchmark_example(path: str | None = None, full_box: bool = False):
    """The reference's viscoelastic attenuation benchmark (EXAMPLES/benchmarks/attenuation/viscoelastic):
    a homogeneous full-space stand-in (rho 2000, vp 3297.849, vs 2222.536 at 18 Hz, Q_kappa 10,
    Q_mu 20, 3 SLS), 37.5 m cubic elements, a vertical unit force at (10, 10, 10) with an 18 Hz Ricker
    STF, one receiver at (510, 510, 510), 1300 steps of 0.5 ms; sources and receivers by z coordinate
    (USE_SOURCES_RECEIVERS_Z).

    The reference meshes [-1500, 1500]^3 (80^3 elements). By default the box here is cut to
    [-787.5, 1312.5]^3 (56^3 elements on the same grid lines): every boundary reflection still reaches
    the receiver after the 0.583 s record ends (shortest reflected P path 2212 m = 0.671 s), and the
    memory variables stay under 4 GiB. Returns (db, sources, receivers, attenuation, fixture dict)."""
    from .attenuation import setup_attenuation
    from .mesh.box import build_box, homogeneous
    z = np.load(path or os.path.join(GOLDEN, "attenuation_benchmark.npz"), allow_pickle=False)
    dt = float(z["DT"])
    h = 37.5
    lo, hi = (-1500.0, 1500.0) if full_box else (-787.5, 1312.5)
    n = int(round((hi - lo) / h))
    db = build_box(n, n, n, (lo, hi, lo, hi, lo, hi), material=homogeneous(2000.0, 3297.849, 2222.536),
                   stacey_faces=())
    db.finalize(dt, stacey=False)
    att = setup_attenuation(db, qmu=20.0, qkappa=10.0, f0_ref=18.0)
    forces = S.read_forcesolution(io.StringIO(str(z["forcesolution"])))
    src = S.setup_force_sources(db, forces, dt, ricker=True, z_coordinates=True)
    st = S.read_stations(io.StringIO(str(z["stations"])))
    rec = S.setup_receivers(db, st, z_coordinates=True)
    return db, src, rec, att, {k: z[k] for k in z.files}


def cpml5_example(path: str | None = None):
    """(db, sources, receivers, ref_seis (nstep, nrec, 3), DT, fixture dict) of the reference example
    CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides: C-PML layers (three elements
    thick) on the four sides and the bottom of a 36 x 36 x 16 CUBIT mesh, free top, a point force
    (1e10 N along (1, -2, -1), Ricker f0 = 0.128 Hz) at 30 km depth, four surface receivers, 2400 steps
    of 0.11 s, f0_FOR_PML = 0.05555 Hz."""
    z = np.load(path or os.path.join(GOLDEN, "cpml5_example.npz"), allow_pickle=False)
    dt = float(z["DT"])
    db = build_from_arrays({k: z[k] for k in z.files}, dt, stacey=False, pml=dict(f0=float(z["f0_FOR_PML"])))
    forces = S.read_forcesolution(io.StringIO(str(z["forcesolution"])))
    src = S.setup_force_sources(db, forces, dt, ricker=True)
    st = S.read_stations(io.StringIO(str(z["stations"])))
    rec = S.setup_receivers(db, st)
    return db, src, rec, z["ref_seis"], dt, {k: z[k] for k in z.files}
