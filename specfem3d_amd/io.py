"""Seismogram and adjoint-source files in the reference's formats.

* ``write_seismograms``: one ASCII file per station and component, ``NET.STA.CHN.sem{d,v,a}``
  (write_seismograms.f90:695), each line ``time value`` as the reference's list-directed FP32
  output (write_output_ASCII_or_binary.f90:103-137). Times are (it-1) DT - t0 (forward) or
  (NSTEP-it) DT - t0 (kernel runs). The channel is a band/instrument code picked from DT
  (band_instrument_code, write_seismograms.f90:1132-1160) plus X/Y/Z (or E/N/Z with the UTM
  projection).
* ``read_adjoint_sources``: ``SEM/NET.STA.CHN.adj`` files (two columns, one line per step,
  compute_arrays_source.f90:203-291) into the (nstep, nrec, 3) array the solvers take.
"""
from __future__ import annotations

import os

import numpy as np


def band_instrument_code(dt: float) -> str:
    """IRIS band + instrument code of synthetic seismograms for a sampling interval dt."""
    if dt >= 1.0:
        return "LX"
    if dt > 0.1:
        return "MX"
    if dt > 0.0125:
        return "BX"
    if dt > 0.004:
        return "HX"
    if dt > 0.001:
        return "CX"
    return "FX"


def channel_names(dt: float, utm: bool = False) -> list[str]:
    b = band_instrument_code(dt)
    return [b + c for c in (("E", "N", "Z") if utm else ("X", "Y", "Z"))]


def fortran_real(x) -> str:
    """A float32 value as the reference's list-directed output prints it: the shortest digits that
    read back to the same float32, fixed notation without a leading zero for 0.1 <= |x| < 1e9
    (``12.``, ``-1.19``, ``-.99``), otherwise scientific with a two-digit signed exponent
    (``2.6315306E-06``), and ``0.`` for zero. (Where two shortest digit strings both read back, the
    nearest is chosen; the reference's runtime sometimes prints the other one.)"""
    v = np.float32(x)
    if v == 0:
        return "-0." if np.signbit(v) else "0."
    a = abs(float(v))
    if 0.1 <= a < 1e9:
        s = np.format_float_positional(v, unique=True, trim="-")
        s = s if "." in s else s + "."
        if s.startswith("0."):
            s = s[1:]
        elif s.startswith("-0."):
            s = "-" + s[2:]
        return s
    s = np.format_float_scientific(v, unique=True, trim="-", exp_digits=2)
    m, e = s.split("e")
    if "." not in m:
        m += "."
    return f"{m}E{e}"


def write_seismograms(out_dir: str, stations, seis: np.ndarray, dt: float, t0: float, kind: str = "d",
                      simulation_type: int = 1, utm: bool = False) -> list[str]:
    """Write (nstep, nrec, 3) seismograms; returns the file paths."""
    os.makedirs(out_dir, exist_ok=True)
    nstep = seis.shape[0]
    it = np.arange(1, nstep + 1, dtype=np.float64)
    if simulation_type == 3:
        tt = (nstep - it) * dt - t0
    else:
        tt = (it - 1) * dt - t0
    times = [fortran_real(t) for t in tt.astype(np.float32)]
    paths = []
    for r, st in enumerate(stations):
        for c, chn in enumerate(channel_names(dt, utm)):
            p = os.path.join(out_dir, f"{st.network}.{st.name}.{chn}.sem{kind}")
            with open(p, "w") as f:
                for t, v in zip(times, seis[:, r, c]):
                    f.write(f" {t} {fortran_real(v)}\n")
            paths.append(p)
    return paths


def read_two_columns(path: str) -> tuple[np.ndarray, np.ndarray]:
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return a[:, 0], a[:, 1]


def read_adjoint_sources(sem_dir: str, stations, nstep: int, dt: float, utm: bool = False) -> np.ndarray:
    """(nstep, nrec, 3) adjoint sources in forward time from ``NET.STA.CHN.adj``; a station without
    files contributes zeros (the reference requires the files of every adjoint station)."""
    out = np.zeros((nstep, len(stations), 3), np.float32)
    for r, st in enumerate(stations):
        for c, chn in enumerate(channel_names(dt, utm)):
            p = os.path.join(sem_dir, f"{st.network}.{st.name}.{chn}.adj")
            if not os.path.exists(p):
                continue
            _, v = read_two_columns(p)
            if v.size < nstep:
                raise ValueError(f"{p}: {v.size} samples, the run has {nstep} steps")
            out[:, r, c] = v[:nstep].astype(np.float32)
    return out


def write_adjoint_sources(sem_dir: str, stations, adj: np.ndarray, dt: float, t0: float, utm: bool = False):
    """Write (nstep, nrec, 3) adjoint sources in forward time as ``NET.STA.CHN.adj`` files."""
    os.makedirs(sem_dir, exist_ok=True)
    nstep = adj.shape[0]
    tt = (np.arange(nstep, dtype=np.float64) * dt - t0).astype(np.float32)
    for r, st in enumerate(stations):
        for c, chn in enumerate(channel_names(dt, utm)):
            with open(os.path.join(sem_dir, f"{st.network}.{st.name}.{chn}.adj"), "w") as f:
                for t, v in zip(tt, adj[:, r, c]):
                    f.write(f" {fortran_real(t)} {fortran_real(v)}\n")
