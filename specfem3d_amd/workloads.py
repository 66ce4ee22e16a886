"""The five workloads the solver loop is measured and checked on (synthetic, seeded).

1. ``homogeneous_box`` -- 20^3-element homogeneous elastic box (10 km cube, vp 2800, vs 1500, rho 2300),
   free top + Stacey on 5 faces, one CMT source at 5 km depth (Ricker, f0 = 1 Hz), DT 0.01 s,
   2000 steps, 10 surface receivers: the reference's CPU-runnable case.
2. ``sea_over_visco_box`` -- 96 x 96 x 48 elements (500 m cubes), top 6 layers acoustic (vp 1500, rho 1020)
   coupled to an elastic half-space with linear depth gradient x (1 + 5% seeded Gaussian random
   field, 2 km correlation), viscoelastic Q_mu = 100 with 3 SLS, one seeded random moment-tensor
   source, 4000 steps, 100 receivers on the sea floor.
3. ``gradient_shots_box`` -- 64 x 64 x 32 elastic box (same gradient + GRF model, no attenuation), 32 shots
   with seeded random locations, moment tensors, Ricker f0 in [0.5, 2] Hz, record lengths in
   [1000, 4000] steps, 50 receivers per shot.
4. ``gradient_adjoint_box`` -- 48^3 elastic box for adjoint sensitivity kernels, 1 source, 4000 steps,
   64 receivers.
5. ``sea_over_visco_large`` -- 160 x 160 x 80 (2.05 M elements, 131.9 M GLL points) viscoelastic with an
   acoustic top layer (10 element layers), 8 shots, 6000 steps.

Each builder accepts a ``scale`` to shrink the element counts for tests; the GLL-point count
``NGLOB = (4 nx + 1)(4 ny + 1)(4 nz + 1)``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import sources as S
from .mesh.box import build_box, gradient_grf_model, homogeneous


@dataclass
class Workload:
    name: str
    db: object
    sources: object
    receivers: object
    nstep: int
    dt: float
    attenuation: object = None
    shots: list = field(default_factory=list)   # batched: list of (SourceSet, ReceiverSet, nstep)
    meta: dict = field(default_factory=dict)


HOMOGENEOUS_BOX_CMT = dict(Mrr=-7.4e19, Mtt=-1.13e21, Mpp=1.21e21, Mrt=3.81e20, Mrp=-8.0e18, Mtp=1.17e21)


def cmt_from_rtp(x, y, depth_km, hdur, Mrr, Mtt, Mpp, Mrt, Mrp, Mtp, tshift=0.0):
    """CMTSOLUTION-style source (dyne-cm) mapped to Cartesian N-m as ``get_cmt``/``locate_source``."""
    s = 1e-7
    return S.CMTSource(x=x, y=y, depth_km=depth_km, tshift=tshift, hdur=hdur,
                       Mxx=Mpp * s, Myy=Mtt * s, Mzz=Mrr * s, Mxy=-Mtp * s, Mxz=Mrp * s, Myz=-Mrt * s)


def homogeneous_box(n: int = 20, nstep: int = 2000, dt: float = 0.01) -> Workload:
    L = 10000.0
    db = build_box(n, n, n, (0.0, L, 0.0, L, -L, 0.0), material=homogeneous(2300.0, 2800.0, 1500.0))
    db.finalize(dt)
    cmt = cmt_from_rtp(5000.0, 5000.0, 5.0, 1.0, **HOMOGENEOUS_BOX_CMT)
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    st = [S.Station(f"X{i + 1:02d}", "DB", 5000.0 + 300.0 * (i % 2), 1000.0 + 900.0 * i, 0.0, 0.0) for i in range(10)]
    rec = S.setup_receivers(db, st)
    return Workload("20x20x20 homogeneous elastic box, 2000 steps", db, src, rec, nstep, dt)


def _model_extent(nx, ny, nz, h):
    return (0.0, nx * h, 0.0, ny * h, -nz * h, 0.0)


def random_cmt(rng, extent, zmin_frac=0.2, zmax_frac=0.8, f0=(1.0, 1.0), z_top=0.0):
    x0, x1, y0, y1, z0, z1 = extent
    x = rng.uniform(x0 + 0.1 * (x1 - x0), x1 - 0.1 * (x1 - x0))
    y = rng.uniform(y0 + 0.1 * (y1 - y0), y1 - 0.1 * (y1 - y0))
    depth = rng.uniform(zmin_frac, zmax_frac) * (z_top - z0)
    M = rng.normal(size=6) * 1e20
    f = rng.uniform(*f0)
    return cmt_from_rtp(x, y, depth / 1000.0, f, *M)


def receiver_grid(extent, n_side, z_below_top=0.0, prefix="R"):
    x0, x1, y0, y1, z0, z1 = extent
    xs = np.linspace(x0 + 0.15 * (x1 - x0), x1 - 0.15 * (x1 - x0), n_side)
    ys = np.linspace(y0 + 0.15 * (y1 - y0), y1 - 0.15 * (y1 - y0), n_side)
    st = []
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            st.append(S.Station(f"{prefix}{i:02d}{j:02d}", "SY", float(y), float(x), 0.0, float(z_below_top)))
    return st


def sea_over_visco_box(scale: float = 1.0, nstep: int = 4000, dt: float = 0.006, seed: int = 2,
                       with_attenuation: bool = True, qmu: float = 100.0, nrec_side: int = 10) -> Workload:
    nx = max(2, int(round(96 * scale)))
    nz = max(8, int(round(48 * scale)))
    h = 500.0
    ext = _model_extent(nx, nx, nz, h)
    nac = 6 if nz >= 12 else 2
    mat = gradient_grf_model(ext, seed=seed, z_top=-nac * h)
    db = build_box(nx, nx, nz, ext, material=mat, acoustic_top_layers=nac,
                   acoustic_material=homogeneous(1020.0, 1500.0, 0.0))
    db.finalize(dt)
    att = None
    if with_attenuation:
        from .attenuation import setup_attenuation
        att = setup_attenuation(db, qmu=qmu, qkappa=9000.0)
    rng = np.random.default_rng(seed)
    seafloor = nac * h
    cmt = random_cmt(rng, ext, zmin_frac=0.3, zmax_frac=0.7, f0=(1.0, 1.0), z_top=0.0)
    cmt.depth_km = max(cmt.depth_km, (seafloor + 2 * h) / 1000.0)
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    rec = S.setup_receivers(db, receiver_grid(ext, nrec_side, z_below_top=seafloor + 1.0))
    return Workload(f"{nx}x{nx}x{nz} acoustic-over-viscoelastic, {nstep} steps", db, src, rec, nstep, dt,
                    attenuation=att, meta=dict(nx=nx, nz=nz, n_acoustic_layers=nac,
                                               qmu=qmu if with_attenuation else None))


def gradient_shots_box(scale: float = 1.0, nshots: int = 32, seed: int = 3, dt: float = 0.006, nrec: int = 50,
                       nstep_range=(1000, 4000)) -> Workload:
    nx = max(2, int(round(64 * scale)))
    nz = max(2, int(round(32 * scale)))
    h = 500.0
    ext = _model_extent(nx, nx, nz, h)
    mat = gradient_grf_model(ext, seed=seed)
    db = build_box(nx, nx, nz, ext, material=mat)
    db.finalize(dt)
    rng = np.random.default_rng(seed)
    shots = []
    for s in range(nshots):
        cmt = random_cmt(rng, ext, f0=(0.5, 2.0))
        src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
        xs = rng.uniform(ext[0], ext[1], nrec)
        ys = rng.uniform(ext[2], ext[3], nrec)
        st = [S.Station(f"S{s:02d}R{r:02d}", "SY", float(ys[r]), float(xs[r]), 0.0, 0.0) for r in range(nrec)]
        rec = S.setup_receivers(db, st)
        n = int(rng.integers(nstep_range[0], nstep_range[1] + 1))
        shots.append((src, rec, n))
    return Workload(f"{nx}x{nx}x{nz} elastic gradient + GRF, {nshots} shots", db, shots[0][0], shots[0][1],
                    max(s[2] for s in shots), dt, shots=shots)


def gradient_adjoint_box(scale: float = 1.0, nstep: int = 4000, dt: float = 0.006, seed: int = 4,
                         nrec_side: int = 8) -> Workload:
    n = max(2, int(round(48 * scale)))
    h = 500.0
    ext = _model_extent(n, n, n, h)
    mat = gradient_grf_model(ext, seed=seed)
    db = build_box(n, n, n, ext, material=mat)
    db.finalize(dt)
    rng = np.random.default_rng(seed)
    cmt = random_cmt(rng, ext, f0=(1.0, 1.0))
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    rec = S.setup_receivers(db, receiver_grid(ext, nrec_side, 0.0))
    return Workload(f"{n}^3 elastic gradient + GRF adjoint, {nstep} steps", db, src, rec, nstep, dt)


def sea_over_visco_large(scale: float = 1.0, nshots: int = 8, nstep: int = 6000, dt: float = 0.006, seed: int = 5,
                         qmu: float = 100.0, nrec_side: int = 8) -> Workload:
    """160 x 160 x 80 elements (scale 1): the top 10 layers acoustic (vp 1500, rho 1020) over the
    gradient + GRF solid with Q_mu = 100 (3 SLS), ``nshots`` seeded random CMT sources below the sea
    floor sharing a sea-floor receiver grid."""
    nx = max(2, int(round(160 * scale)))
    nz = max(8, int(round(80 * scale)))
    h = 500.0
    ext = _model_extent(nx, nx, nz, h)
    nac = max(1, nz // 8)
    mat = gradient_grf_model(ext, seed=seed, z_top=-nac * h)
    db = build_box(nx, nx, nz, ext, material=mat, acoustic_top_layers=nac,
                   acoustic_material=homogeneous(1020.0, 1500.0, 0.0))
    db.finalize(dt)
    from .attenuation import setup_attenuation
    att = setup_attenuation(db, qmu=qmu, qkappa=9000.0)
    rng = np.random.default_rng(seed)
    seafloor = nac * h
    rec = S.setup_receivers(db, receiver_grid(ext, nrec_side, z_below_top=seafloor + 1.0))
    shots = []
    for _ in range(nshots):
        cmt = random_cmt(rng, ext, zmin_frac=0.3, zmax_frac=0.7, f0=(0.5, 2.0), z_top=0.0)
        cmt.depth_km = max(cmt.depth_km, (seafloor + 2 * h) / 1000.0)
        shots.append((S.setup_cmt_sources(db, [cmt], dt, ricker=True), rec, nstep))
    return Workload(f"{nx}x{nx}x{nz} acoustic-over-viscoelastic, {nshots} shots, {nstep} steps", db,
                    shots[0][0], rec, nstep, dt, attenuation=att, shots=shots,
                    meta=dict(nx=nx, nz=nz, n_acoustic_layers=nac))


def gaussian_bump(db, center=None, sigma=None) -> np.ndarray:
    """Smooth interior Gaussian exp(-|x - c|^2 / (2 sigma^2)) at every local point, (nspec, 125);
    defaults: c = centre of the mesh, sigma = 1/8 of its smallest side."""
    ib = db.ibool.reshape(db.nspec, 125)
    x, y, z = db.xstore[ib].astype(np.float64), db.ystore[ib].astype(np.float64), db.zstore[ib].astype(np.float64)
    lo = np.array([db.xstore.min(), db.ystore.min(), db.zstore.min()], np.float64)
    hi = np.array([db.xstore.max(), db.ystore.max(), db.zstore.max()], np.float64)
    c = (lo + hi) / 2 if center is None else np.asarray(center, np.float64)
    s = (hi - lo).min() / 8 if sigma is None else float(sigma)
    return np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))


def nglob_box(nx, ny, nz) -> int:
    return (4 * nx + 1) * (4 * ny + 1) * (4 * nz + 1)
