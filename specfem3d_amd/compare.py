"""Seismogram comparison metrics, as the reference's check script reports them
(utils/scripts/compare_seismogram_correlations.py:46-105 time shift, :236-317 per-trace metrics).

Per trace (reference ``ref``, result ``syn``, sample interval ``dt``):

* ``err``: L2 misfit normalised by the reference, ``|ref - syn| / |ref|`` (unnormalised when the
  reference trace is zero);
* ``corr``: the smallest entry of the 2x2 correlation-coefficient matrix of (ref, syn) when both
  traces are non-zero; the variance of ``ref - syn`` when exactly one is zero; 1 for two zero
  traces;
* ``shift``: time lag of the cross-correlation maximum of ref against syn, refined by the vertex of
  the parabola through the maximum and its two neighbours (cyclic at the ends); 0 for a zero
  reference.

``compare`` reduces them over traces as the script does (smallest correlation, largest error,
largest absolute shift) and adds the aggregate rel-L2 over all traces.
"""
from __future__ import annotations

import numpy as np


def xcorr_time_shift(x: np.ndarray, y: np.ndarray, dt: float) -> float:
    """Lag (in time units) at which the cross-correlation of x with y peaks, with sub-sample
    parabolic refinement; positive when x leads y."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    if x.shape != y.shape:
        raise ValueError("traces of different lengths")
    n = x.size
    cc = np.correlate(x, y, mode="full")        # lags -(n-1) .. n-1
    m = cc.size
    k = int(np.argmax(cc))
    lag = float(k + 1 - n)
    left = cc[k - 1] if k >= 1 else cc[m - 1]
    right = cc[k + 1] if k < m - 1 else cc[0]
    den = left - 2.0 * cc[k] + right
    if den != 0.0:
        lag += 0.5 * (left - right) / den
    return lag * dt


def trace_metrics(ref: np.ndarray, syn: np.ndarray, dt: float) -> tuple[float, float, float]:
    """(corr, err, shift) of one trace pair."""
    ref = np.asarray(ref, np.float64)
    syn = np.asarray(syn, np.float64)
    nref, nsyn, ndiff = np.linalg.norm(ref), np.linalg.norm(syn), np.linalg.norm(ref - syn)
    err = ndiff / nref if nref > 0.0 else ndiff
    if nref > 0.0 and nsyn > 0.0:
        corr = float(np.min(np.corrcoef(ref, syn)))
    elif ndiff > 0.0:
        corr = float(np.min(np.cov(ref - syn)))
    else:
        corr = 1.0
    shift = xcorr_time_shift(ref, syn, dt) if nref > 0.0 else 0.0
    return corr, float(err), float(shift)


def compare(ref: np.ndarray, syn: np.ndarray, dt: float) -> dict:
    """Metrics over seismograms shaped (nstep, ntraces[, ncomp]): per-trace rows and the reductions."""
    ref = np.asarray(ref)
    syn = np.asarray(syn)
    if ref.shape != syn.shape:
        raise ValueError("seismogram arrays of different shapes")
    r2 = ref.reshape(ref.shape[0], -1)
    s2 = syn.reshape(syn.shape[0], -1)
    rows = [trace_metrics(r2[:, t], s2[:, t], dt) for t in range(r2.shape[1])]
    corr = np.array([r[0] for r in rows])
    err = np.array([r[1] for r in rows])
    shift = np.array([r[2] for r in rows])
    den = float(np.linalg.norm(r2.astype(np.float64)))
    rel_l2 = float(np.linalg.norm((r2 - s2).astype(np.float64)) / den) if den > 0 else float("nan")
    return {"traces": rows, "corr_min": float(corr.min()), "err_max": float(err.max()),
            "shift_max": float(shift[np.argmax(np.abs(shift))]), "rel_l2": rel_l2}
