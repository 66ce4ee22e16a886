"""Oracle pin of the time-reversed reconstruction (the reference's classic SIMULATION_TYPE=3,
iterate_time.F90:233-290): the backward field restarts from the forward run's last frame and steps
with b_deltat = -DT, subtracting the forward run's Stacey records and adding the sources.

The explicit Newmark scheme is reversible: undoing forward step n needs that step's own Stacey record
and source time. With those (one step earlier than the reference's indices), the reconstruction must
retrace every forward frame to rounding (FP64) with and without Stacey faces -- a wrong sign of
b_deltat, a wrong replay sign or a misplaced source fails it. With the reference's own indices (record
of forward step NSTEP-it+1, source time (NSTEP-it) DT - t0) the retrace is off by O(dt): that is the
reference's behaviour, which the GPU path reproduces (tests/test_gpu_reversal.py).
"""
import numpy as np
import pytest

from conftest import small_elastic_box


def forward_with_records(oracle_mod, db, src, rec, n, prec):
    f = oracle_mod.OracleSolver(db, src, rec, prec)
    f.stf_table = src.stf_table(n)
    recs, disp = [], []
    for _ in range(n):
        f.absorb_record = np.zeros(3 * 25 * max(db.abs_faces.n, 1), f.dt_np)
        f.step()
        recs.append(f.absorb_record.copy())
        disp.append(f.displ.copy())
    return recs, disp, (f.displ, f.veloc, f.accel)


@pytest.mark.parametrize("stacey", [False, True])
def test_exact_reversal_retraces_forward_frames(oracle_mod, stacey):
    db, src, rec = small_elastic_box(n=4, dt=0.004, nrec=4, stacey=stacey)
    n = 300
    recs, disp, last = forward_with_records(oracle_mod, db, src, rec, n, "f64")
    b = oracle_mod.reversed_solver(db, src, rec, n, last, "f64")
    worst = 0.0
    for it in range(1, n - 10):
        b.backward_step(recs[n - it - 1], n - it - 1)       # undo forward step n - it
        ref = disp[n - it - 1]
        worst = max(worst, float(np.linalg.norm(b.displ - ref) / np.linalg.norm(ref)))
    print(f"stacey={stacey}: worst retrace rel-L2 {worst:.2e}")
    assert worst < 1e-9


def test_reference_indices_are_one_step_late(oracle_mod):
    db, src, rec = small_elastic_box(n=4, dt=0.004, nrec=4, stacey=True)
    n = 300
    recs, disp, last = forward_with_records(oracle_mod, db, src, rec, n, "f64")
    b = oracle_mod.reversed_solver(db, src, rec, n, last, "f64")
    for it in range(1, 51):
        b.backward_step(recs[n - it], n - it)
    e = float(np.linalg.norm(b.displ - disp[n - 51]) / np.linalg.norm(disp[n - 51]))
    print(f"reference indices, 50 backward steps: rel-L2 {e:.3e}")
    assert 1e-3 < e < 1.0
