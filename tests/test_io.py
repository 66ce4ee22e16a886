"""Seismogram / adjoint-source files against the reference's own output (20^3 homogeneous box, FP32 CPU solver).

The head of the reference's ``OUTPUT_FILES/DB.X01.HXX.semd`` is committed in tests/golden. Writing
the reference's FP32 seismogram values (tests/golden/homogeneous_box_reference_seismograms.npz) must reproduce
its file name, its time column text exactly, and values that read back to the same float32 (the
reference runtime's choice between two equally short digit strings is not reproduced everywhere).
"""
import os

import numpy as np

from conftest import GOLDEN


def test_semd_files_match_reference_output(tmp_path):
    from specfem3d_amd import io
    from specfem3d_amd.sources import Station
    ref = np.load(os.path.join(GOLDEN, "homogeneous_box_reference_seismograms.npz"))
    seis = ref["fp32"]
    stations = [Station(f"X{i + 1:02d}", "DB", 0.0, 0.0, 0.0, 0.0) for i in range(seis.shape[1])]
    t0 = 1.2                         # Ricker f0 = 1 Hz: t0 = 1.2 / f0 (times are (it-1) DT - t0 in FP64)
    paths = io.write_seismograms(str(tmp_path), stations, seis, 0.01, t0)
    assert os.path.basename(paths[0]) == "DB.X01.HXX.semd"
    assert len(paths) == 30
    mine = open(paths[0]).read().splitlines()
    theirs = open(os.path.join(GOLDEN, "DB.X01.HXX.semd.head")).read().splitlines()
    assert len(mine) == seis.shape[0]
    for a, b in zip(mine, theirs):
        ta, va = a.split()
        tb, vb = b.split()
        assert ta == tb
        assert np.float32(float(va)) == np.float32(float(vb))
    exact = sum(a == b for a, b in zip(mine, theirs))
    assert exact >= 0.8 * len(theirs)


def test_fortran_real_formatting():
    from specfem3d_amd.io import fortran_real
    assert fortran_real(0.0) == "0."
    assert fortran_real(-1.2) == "-1.2"
    assert fortran_real(np.float32(-0.99)) == "-.99"
    assert fortran_real(12.0) == "12."
    assert fortran_real(np.float32(2.6315306e-06)) == "2.6315306E-06"
    assert fortran_real(np.float32(1.4716e-07)) == "1.4716E-07"
    for v in np.random.default_rng(0).normal(size=200).astype(np.float32) * np.float32(1e-5):
        assert np.float32(float(fortran_real(v))) == v


def test_adjoint_source_files_roundtrip(tmp_path):
    from specfem3d_amd import io
    from specfem3d_amd.sources import Station
    st = [Station("A1", "XX", 0, 0, 0, 0), Station("B2", "YY", 0, 0, 0, 0)]
    adj = np.random.default_rng(1).normal(size=(50, 2, 3)).astype(np.float32)
    io.write_adjoint_sources(str(tmp_path), st, adj, 0.005, 1.2)
    assert os.path.exists(tmp_path / "XX.A1.HXZ.adj")
    back = io.read_adjoint_sources(str(tmp_path), st, 50, 0.005)
    assert np.array_equal(back, adj)


def test_band_codes():
    from specfem3d_amd.io import band_instrument_code
    assert [band_instrument_code(d) for d in (2.0, 0.5, 0.05, 0.01, 0.002, 0.0005)] == ["LX", "MX", "BX", "HX", "CX", "FX"]
