"""Time-reversed reconstruction (the reference's classic SIMULATION_TYPE=3, iterate_time.F90:233-290)
through the C-ABI: a forward run with SAVE_FORWARD stores the Stacey record of every step and its
last frame; the backward field restarts from that frame with b_deltat = -DT and replays the
records in reverse.

* with the reference's indices (record of forward step NSTEP-it+1, sources at (NSTEP-it) DT - t0)
  the backward field equals the oracle's backward field bit for bit (exact kernels, ordered
  assembly) at every checked step -- including the O(dt) drift those indices cause;
* with the exactly reversing indices (one step earlier) it retraces the forward displacement to
  FP32 rounding.
"""
import numpy as np
import pytest

from conftest import small_elastic_box

pytestmark = pytest.mark.gpu

N = 300


@pytest.fixture(scope="module")
def forward(oracle_mod):
    from specfem3d_amd.adjoint import ReversalKernelSolver
    db, src, rec = small_elastic_box(n=4, dt=0.004, nrec=4)
    print(f"[gpu] reversal: forward {N} steps with records", flush=True)
    seis, recs, last, disp = ReversalKernelSolver.record_forward(db, src, rec, N, exact=True, assembly=1,
                                                                 keep_displ=True)
    return dict(db=db, src=src, rec=rec, recs=recs, last=last, disp=disp)


def test_forward_records_equal_oracle_records(forward, oracle_mod):
    f = forward
    o = oracle_mod.OracleSolver(f["db"], f["src"], f["rec"], "f32")
    o.stf_table = f["src"].stf_table(N)
    for it in range(N):
        o.absorb_record = np.zeros(f["recs"].shape[1], np.float32)
        o.step()
        if it % 50 == 0 or it == N - 1:
            assert np.array_equal(o.absorb_record, f["recs"][it]), it
            assert np.array_equal(o.displ, f["disp"][it]), it


def test_backward_field_equals_oracle_reversal(forward, oracle_mod):
    from specfem3d_amd.adjoint import ReversalKernelSolver
    f = forward
    rs = ReversalKernelSolver(f["db"], f["src"], f["rec"], N, exact=True, assembly=1)
    rs.start_from(f["last"])
    ob = oracle_mod.reversed_solver(f["db"], f["src"], f["rec"], N, f["last"], "f32")
    for it in range(1, 251):
        rs.backward_step(f["recs"])
        ob.backward_step(f["recs"][N - it], N - it)
        if it % 50 == 0:
            b = rs.b_displ()
            drift = float(np.linalg.norm(b - f["disp"][N - it - 1]) / np.linalg.norm(f["disp"][N - it - 1]))
            print(f"[gpu] backward step {it}: equal to oracle {np.array_equal(b, ob.displ)}, "
                  f"drift from forward frame {drift:.3e}", flush=True)
            assert np.array_equal(b, ob.displ), it
    rs.close()


def test_exact_reversal_retraces_forward_frames(forward):
    from specfem3d_amd.adjoint import ReversalKernelSolver
    f = forward
    rs = ReversalKernelSolver(f["db"], f["src"], f["rec"], N, exact=True, assembly=1)
    rs.start_from(f["last"])
    worst = 0.0
    for it in range(1, 251):
        rs.backward_step(f["recs"], exact_reversal=True)
        if it % 25 == 0:
            ref = f["disp"][N - it - 1]
            e = float(np.linalg.norm(rs.b_displ() - ref) / np.linalg.norm(ref))
            worst = max(worst, e)
    rs.close()
    print(f"[gpu] exact reversal: worst retrace rel-L2 over 250 steps {worst:.3e}", flush=True)
    assert worst < 1e-4
