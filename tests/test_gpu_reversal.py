"""Time-reversed reconstruction (the reference's classic SIMULATION_TYPE=3, iterate_time.F90:233-290)
through the C-ABI: a forward run with SAVE_FORWARD stores the Stacey record of every step and its
last frame; the backward field restarts from that frame with b_deltat = -DT and replays the
records in reverse (record NSTEP-it+1, sources at (NSTEP-it) DT - t0, as the reference indexes
them). Its displacement must retrace the forward displacement."""
import numpy as np
import pytest

from conftest import small_elastic_box

pytestmark = pytest.mark.gpu


def test_backward_field_retraces_forward_run():
    from specfem3d_amd.adjoint import ReversalKernelSolver
    db, src, rec = small_elastic_box(n=4, dt=0.004, nrec=4)
    n = 300
    print(f"[gpu] reversal: forward {n} steps with records", flush=True)
    seis, recs, last, disp = ReversalKernelSolver.record_forward(db, src, rec, n, exact=True, assembly=1,
                                                                 keep_displ=True)
    rs = ReversalKernelSolver(db, src, rec, n, exact=True, assembly=1)
    rs.start_from(last)
    errs = {}
    for it in range(1, n + 1):
        rs.backward_step(recs)
        if it % 50 == 0 or it == n - 1:
            b = rs.b_displ()
            # the reference pairs backward step `it` with the forward frame NSTEP - it + 1 (records)
            for off in (0, 1):
                k = n - it - off          # 0-based index of forward step n - it - off + 1
                if 0 <= k < n:
                    ref = disp[k]
                    errs[(it, off)] = float(np.linalg.norm(b - ref) / max(np.linalg.norm(ref), 1e-30))
    rs.close()
    for key in sorted(errs):
        print(f"[gpu] backward step {key[0]}, forward frame offset {key[1]}: rel-L2 {errs[key]:.3e}", flush=True)
    best = {it: min(v for (i, o), v in errs.items() if i == it) for it in {i for i, _ in errs}}
    assert max(best.values()) < 1e-3
