"""Seismogram comparison metrics (specfem3d_amd/compare.py) against what their definitions fix:
identical traces, scaled traces, known shifts of a smooth pulse, zero traces."""
import numpy as np

from specfem3d_amd.compare import compare, trace_metrics, xcorr_time_shift


def ricker(t, f0=1.0, t0=1.5):
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def test_identical_and_scaled_traces():
    t = np.arange(800) * 0.01
    x = ricker(t)
    assert trace_metrics(x, x, 0.01) == (1.0, 0.0, 0.0)
    c, e, s = trace_metrics(x, 2 * x, 0.01)
    assert abs(c - 1.0) < 1e-12 and abs(e - 1.0) < 1e-12 and abs(s) < 1e-12


def test_shift_of_a_delayed_pulse_is_recovered_to_a_fraction_of_a_sample():
    dt = 0.01
    t = np.arange(1000) * dt
    for d in (0.0, 0.034, -0.05, 0.12):
        ref = ricker(t)
        syn = ricker(t, t0=1.5 + d)             # syn lags ref by d
        s = xcorr_time_shift(ref, syn, dt)
        assert abs(s + d) < 0.05 * dt, (d, s)   # ref leads: negative lag of -d
        assert trace_metrics(ref, syn, dt)[2] == s


def test_zero_traces():
    z = np.zeros(100)
    x = np.sin(np.arange(100) * 0.1)
    assert trace_metrics(z, z, 0.1) == (1.0, 0.0, 0.0)
    c, e, s = trace_metrics(z, x, 0.1)
    assert abs(e - np.linalg.norm(x)) < 1e-12 and s == 0.0 and abs(c - np.var(x, ddof=1)) < 1e-12


def test_reductions_over_traces():
    dt = 0.01
    t = np.arange(600) * dt
    ref = np.stack([ricker(t), ricker(t, t0=2.0), ricker(t, t0=2.5)], 1)[:, :, None].repeat(3, 2)
    syn = ref.copy()
    syn[:, 1, 2] *= 1.1
    syn[:, 2, 0] = ricker(t, t0=2.5 + 0.03)
    out = compare(ref, syn, dt)
    assert len(out["traces"]) == 9
    assert abs(out["err_max"] - max(0.1, np.linalg.norm(ref[:, 2, 0] - syn[:, 2, 0]) / np.linalg.norm(ref[:, 2, 0]))) < 1e-12
    assert abs(out["shift_max"] + 0.03) < 0.05 * dt
    assert out["corr_min"] < 1.0
    want = np.linalg.norm(ref - syn) / np.linalg.norm(ref)
    assert abs(out["rel_l2"] - want) < 1e-12
