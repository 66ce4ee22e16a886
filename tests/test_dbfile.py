"""Solver database files (proc000000_external_mesh.bin, proc000000_attenuation.bin) and the
generator's arrays against closed forms.

* round trips: elastic (warped, irregular), coupled fluid-solid with Stacey, fluid-only; every array the
  solver uses comes back bit for bit, and the combined inverse mass matrices equal finalize()'s;
* regular elements: a box's cubes are regular with xix = 2/h and det = (h/2)^3, a warped box's interior
  elements are irregular (calc_jacobian.f90:184-345);
* mass matrices of a homogeneous box: the elastic mass sums to rho V, the acoustic one to V / kappa;
* Stacey mass terms (define_mass_matrices.f90:353-522): summed over the points, the elastic x term is
  dt/2 [rho vp (area of the x faces) + rho vs (area of the y and bottom faces)] (z likewise), and the
  acoustic term dt/2 (absorbing area) / (rho c). A wrong normal projection, a lost face weight or a
  swapped vp/vs fails these.
"""
import os

import numpy as np
import pytest

from conftest import small_coupled_box, small_elastic_box, smooth_warp


def _roundtrip(db, tmp_path):
    from specfem3d_amd.mesh.dbfile import read_solver_database, write_solver_database
    p = os.path.join(tmp_path, "proc000000_external_mesh.bin")
    write_solver_database(db, p)
    return read_solver_database(p, db.dt, stacey=db.stacey)


ARRAYS = ("ibool", "xstore", "ystore", "zstore", "irregular_element_number", "jac9", "jacobian", "kappastore",
          "mustore", "rhostore", "rho_vp", "rho_vs", "ispec_is_elastic", "ispec_is_acoustic", "rmassx", "rmassy",
          "rmassz", "rmass_acoustic")


@pytest.mark.parametrize("case", ["elastic_warped", "coupled", "fluid"])
def test_database_file_roundtrip(tmp_path, case):
    from specfem3d_amd.mesh.box import build_box, homogeneous
    if case == "elastic_warped":
        db, _, _ = small_elastic_box(n=3, warp=smooth_warp(1500.0))
        assert db.nspec_irregular > 0
    elif case == "coupled":
        db, _, _ = small_coupled_box(n=4)
        assert db.coupling_faces.n > 0
    else:
        db = build_box(3, 3, 2, (0, 900.0, 0, 900.0, -600.0, 0), material=homogeneous(1020.0, 1500.0, 0.0),
                       acoustic_top_layers=2, acoustic_material=homogeneous(1020.0, 1500.0, 0.0))
        db.finalize(0.002)
    db2 = _roundtrip(db, str(tmp_path))
    for k in ARRAYS:
        a, b = np.asarray(getattr(db, k)), np.asarray(getattr(db2, k))
        if case == "fluid" and k in ("rho_vp", "rho_vs"):
            continue
        assert a.shape == b.shape and np.array_equal(a, b), k
    for f in ("abs_faces", "free_faces", "coupling_faces"):
        a, b = getattr(db, f), getattr(db2, f)
        for x in ("ispec", "ijk", "normal", "jacobian2Dw"):
            assert np.array_equal(getattr(a, x), getattr(b, x)), (f, x)
    assert (db2.xix_regular, db2.jacobian_regular, db2.nspec_irregular) == (db.xix_regular, db.jacobian_regular,
                                                                              db.nspec_irregular)


def test_attenuation_file_roundtrip(tmp_path):
    from specfem3d_amd.attenuation import setup_attenuation
    from specfem3d_amd.mesh.dbfile import read_attenuation_bin, write_attenuation_bin
    db, _, _ = small_coupled_box(n=3)
    att = setup_attenuation(db, qmu=80.0, qkappa=400.0)
    p = os.path.join(str(tmp_path), "proc000000_attenuation.bin")
    write_attenuation_bin(att, db.nspec, p)
    fc, sc, fck, sck = read_attenuation_bin(p)
    assert np.array_equal(fc, att.factor_common) and np.array_equal(fck, att.factor_common_kappa)
    assert np.array_equal(sc, att.scale_factor) and np.array_equal(sck, att.scale_factor_kappa)
    el = db.ispec_is_elastic
    assert np.array_equal((db.mustore[el] * sc[el]).astype(np.float32), att.mustore[el])


def test_regular_classification_closed_form():
    from specfem3d_amd.mesh.box import build_box, homogeneous
    h = 250.0
    db = build_box(3, 3, 3, (0, 3 * h, 0, 3 * h, -3 * h, 0), material=homogeneous(2000.0, 3000.0, 1700.0))
    assert db.nspec_irregular == 0 and (db.irregular_element_number == 0).all()
    assert db.xix_regular == np.float32(2.0 / h) and db.jacobian_regular == np.float32((h / 2) ** 3)
    w = build_box(3, 3, 3, (0, 3 * h, 0, 3 * h, -3 * h, 0), material=homogeneous(2000.0, 3000.0, 1700.0),
                  warp=smooth_warp(3 * h))
    # the warp moves every interior node; the centre element's 8 corners all move by the same vector (the
    # warp is symmetric about the box centre), so it stays a cube of edge h and is the one regular element
    assert w.nspec_irregular == 26 and w.irregular_element_number[13] == 0
    assert sorted(w.irregular_element_number[np.arange(27) != 13]) == list(range(1, 27))
    assert abs(w.xix_regular * h / 2 - 1) < 1e-5 and np.all(w.jacobian > 0)   # edge from moved float32 anchors


def test_mass_and_stacey_terms_closed_form():
    from specfem3d_amd.mesh.box import build_box, homogeneous
    rho, vp, vs, dt = 2100.0, 3200.0, 1800.0, 0.003
    Lx, Ly, Lz = 1200.0, 900.0, 600.0
    db = build_box(4, 3, 2, (0, Lx, 0, Ly, -Lz, 0), material=homogeneous(rho, vp, vs))
    db.finalize(dt)
    p = db.mass_parts
    V = Lx * Ly * Lz
    assert abs(p["rmass"].astype(np.float64).sum() / (rho * V) - 1) < 1e-6
    ax, ay, ab = 2 * Ly * Lz, 2 * Lx * Lz, Lx * Ly        # x faces, y faces, bottom
    want = {"rmassx": 0.5 * dt * (rho * vp * ax + rho * vs * (ay + ab)),
            "rmassy": 0.5 * dt * (rho * vp * ay + rho * vs * (ax + ab)),
            "rmassz": 0.5 * dt * (rho * vp * ab + rho * vs * (ax + ay))}
    for k, v in want.items():
        assert abs(p[k].astype(np.float64).sum() / v - 1) < 1e-6, k
    f = build_box(4, 3, 2, (0, Lx, 0, Ly, -Lz, 0), material=homogeneous(1020.0, 1500.0, 0.0), acoustic_top_layers=2,
                  acoustic_material=homogeneous(1020.0, 1500.0, 0.0))
    f.finalize(dt)
    q = f.mass_parts
    kappa = 1020.0 * 1500.0 ** 2
    assert abs(q["rmass_acoustic"].astype(np.float64).sum() / (V / kappa) - 1) < 1e-6
    assert abs(q["rmassz_acoustic"].astype(np.float64).sum() / (0.5 * dt * (ax + ay + ab) / (1020.0 * 1500.0)) - 1) < 1e-6


def test_database_file_roundtrip_with_cpml(tmp_path):
    """C-PML records (save_arrays_solver.F90:186-218): widths, separation distance, regions, element lists
    and damping profiles come back bit for bit, the absorbing faces are kept without Stacey mass terms,
    and the convolution coefficients recomputed from them equal the original ones."""
    from conftest import small_pml_box
    from specfem3d_amd.mesh.dbfile import read_solver_database, write_solver_database
    db, _, _, _ = small_pml_box(n=5, layers=2, warp=None)
    p = str(tmp_path / "proc000000_external_mesh.bin")
    write_solver_database(db, p)
    rd = read_solver_database(p, db.dt, stacey=False, pml=True)
    a, b = db.pml, rd.pml
    assert np.array_equal(a.cpml_to_spec, b.cpml_to_spec) and np.array_equal(a.regions, b.regions)
    assert np.array_equal(a.width, b.width) and a.min_distance == b.min_distance
    for x, y in ((a.d, b.d), (a.K, b.K), (a.alpha, b.alpha), (a.coef_alpha, b.coef_alpha),
                 (a.coef_beta, b.coef_beta), (a.coef_abar, b.coef_abar), (a.coef_strain, b.coef_strain)):
        assert np.array_equal(x, y)
    assert rd.abs_faces.n == db.abs_faces.n and not rd.stacey
    assert np.array_equal(rd.rmassx, db.rmassx)
