import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build_oracle()
    return oracle


def small_elastic_box(n=4, warp=None, stacey=True, dt=0.004, nrec=6, seed=0):
    """Small homogeneous/gradient box with one CMT source and a few receivers (test helper)."""
    from specfem3d_amd.mesh.box import build_box, gradient_grf_model
    from specfem3d_amd import sources as S
    L = 500.0 * n
    ext = (0.0, L, 0.0, L, -L, 0.0)
    faces = (1, 2, 3, 4, 5) if stacey else ()
    db = build_box(n, n, n, ext, material=gradient_grf_model(ext, seed=seed, corr_len=700.0), stacey_faces=faces,
                   warp=warp)
    db.finalize(dt, stacey=stacey)
    rng = np.random.default_rng(seed)
    cmt = S.CMTSource(x=float(rng.uniform(0.3, 0.7) * L), y=float(rng.uniform(0.3, 0.7) * L),
                      depth_km=float(rng.uniform(0.3, 0.7) * L / 1000.0), tshift=0.0, hdur=4.0,
                      Mxx=1e13, Myy=-2e13, Mzz=0.7e13, Mxy=0.3e13, Mxz=-0.5e13, Myz=0.2e13)
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    st = [S.Station(f"R{i}", "T", float(rng.uniform(0.1, 0.9) * L), float(rng.uniform(0.1, 0.9) * L), 0.0,
                    float(rng.uniform(0.0, 0.5) * L)) for i in range(nrec)]
    rec = S.setup_receivers(db, st)
    return db, src, rec


def small_coupled_box(n=4, dt=0.004, nrec=6, seed=0, warp=None):
    """Box with a fluid top layer (free surface on top) over a gradient solid, Stacey on the sides and
    bottom, one CMT source and receivers in the solid (test helper)."""
    from specfem3d_amd.mesh.box import build_box, gradient_grf_model, homogeneous
    from specfem3d_amd import sources as S
    L = 500.0 * n
    ext = (0.0, L, 0.0, L, -L, 0.0)
    db = build_box(n, n, n, ext, material=gradient_grf_model(ext, seed=seed, corr_len=700.0),
                   acoustic_top_layers=1, acoustic_material=homogeneous(1020.0, 1500.0, 0.0),
                   stacey_faces=(1, 2, 3, 4, 5), warp=warp)
    db.finalize(dt, stacey=True)
    rng = np.random.default_rng(seed)
    cmt = S.CMTSource(x=float(rng.uniform(0.3, 0.7) * L), y=float(rng.uniform(0.3, 0.7) * L),
                      depth_km=float(rng.uniform(0.45, 0.75) * L / 1000.0), tshift=0.0, hdur=4.0,
                      Mxx=1e13, Myy=-2e13, Mzz=0.7e13, Mxy=0.3e13, Mxz=-0.5e13, Myz=0.2e13)
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    st = [S.Station(f"R{i}", "T", float(rng.uniform(0.1, 0.9) * L), float(rng.uniform(0.1, 0.9) * L), 0.0,
                    float(rng.uniform(0.35, 0.6) * L)) for i in range(nrec)]
    rec = S.setup_receivers(db, st)
    assert db.ispec_is_elastic[src.ispec].all() and db.ispec_is_elastic[rec.ispec].all()
    return db, src, rec


def smooth_warp(L, amp=0.06):
    """Smooth interior displacement of the anchors (keeps the box faces planar): irregular elements."""
    def w(x, y, z):
        sx = np.sin(np.pi * x / L)
        sy = np.sin(np.pi * y / L)
        sz = np.sin(np.pi * (-z) / L)
        d = amp * L * sx * sy * sz
        return x + 0.5 * d, y - 0.3 * d, z + 0.4 * d
    return w


def small_pml_box(n=6, layers=2, warp=None, dt=0.004, nrec=6, seed=0, attenuation=False, top=False):
    """Gradient box with C-PML layers ``layers`` elements thick on the four sides and the bottom (free
    top; with ``top`` a layer replaces the free surface too, PML_INSTEAD_OF_FREE_SURFACE), one CMT
    source and receivers inside the layers' interior (test helper)."""
    from specfem3d_amd.mesh.box import build_box, gradient_grf_model
    from specfem3d_amd import pml as P, sources as S
    L = 500.0 * n
    ext = (0.0, L, 0.0, L, -L, 0.0)
    db = build_box(n, n, n, ext, material=gradient_grf_model(ext, seed=seed, corr_len=700.0),
                   stacey_faces=(1, 2, 3, 4, 5, 6) if top else (1, 2, 3, 4, 5), warp=warp)
    faces = ("xmin", "xmax", "ymin", "ymax", "bottom") + (("top",) if top else ())
    cs, rg = P.box_regions(db, layers, faces)
    db.pml = P.set_damping(db, cs, rg, f0=4.0, instead_of_free_surface=top)
    db.finalize(dt, stacey=False)
    att = None
    if attenuation:
        from specfem3d_amd.attenuation import setup_attenuation
        att = setup_attenuation(db, qmu=60.0, qkappa=200.0, f0_ref=4.0)
    rng = np.random.default_rng(seed)
    zs = (0.42, 0.58) if top else (0.3, 0.5)
    zr = (0.4, 0.6) if top else (0.0, 0.3)
    cmt = S.CMTSource(x=float(rng.uniform(0.42, 0.58) * L), y=float(rng.uniform(0.42, 0.58) * L),
                      depth_km=float(rng.uniform(*zs) * L / 1000.0), tshift=0.0, hdur=4.0,
                      Mxx=1e13, Myy=-2e13, Mzz=0.7e13, Mxy=0.3e13, Mxz=-0.5e13, Myz=0.2e13)
    src = S.setup_cmt_sources(db, [cmt], dt, ricker=True)
    st = [S.Station(f"R{i}", "T", float(rng.uniform(0.1, 0.9) * L), float(rng.uniform(0.1, 0.9) * L), 0.0,
                    float(rng.uniform(*zr) * L)) for i in range(nrec)]
    rec = S.setup_receivers(db, st)
    return db, src, rec, att
