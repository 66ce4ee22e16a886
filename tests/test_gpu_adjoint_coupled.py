"""GPU checkpointed kernel run on a coupled fluid-solid mesh (KernelSolver through the C-ABI) against
the oracle: the fluid kernels (compute_kernels_acoustic_cuda) in both pairings -- the reference's
(adjoint potential) and the misfit-gradient one (adjoint potential_dot_dot) -- and the solid
kernels of the same run.

Targets: every kernel within 1e-4 aggregate rel-L2 of the FP64 oracle; the reconstructed forward
seismograms within 1e-5 of the FP32 oracle (bit for bit in exact/ordered mode).
"""
import copy

import numpy as np
import pytest

from conftest import rel_l2, small_coupled_box

pytestmark = pytest.mark.gpu

NSTEP = 400


@pytest.fixture(scope="module")
def problem(oracle_mod):
    from specfem3d_amd import adjoint as A
    db, src, rec = small_coupled_box(n=4, dt=0.004, nrec=6, seed=1)
    L = 2000.0
    ib = db.ibool.reshape(db.nspec, 125)
    xl, yl, zl = db.xstore[ib], db.ystore[ib], db.zstore[ib]
    s = L / 6
    phi = np.exp(-((xl - L / 2) ** 2 + (yl - L / 2) ** 2 + (zl + 250.0) ** 2) / (2 * s * s))
    d = copy.copy(db)
    d.kappastore = (db.kappastore.astype(np.float64) * (1 + 0.03 * phi * db.ispec_is_acoustic[:, None])
                    ).astype(np.float32).reshape(db.kappastore.shape)
    d.finalize(db.dt)
    obs = oracle_mod.OracleSolver(d, src, rec, "f32").run(NSTEP)
    syn = oracle_mod.OracleSolver(db, src, rec, "f32").run(NSTEP)
    adj = A.waveform_adjoint_source(syn, obs)
    ref = {}
    for field in ("potential", "potential_dot_dot"):
        for prec in ("f32", "f64"):
            out = oracle_mod.kernel_run(db, src, rec, NSTEP, adj.astype(np.float64 if prec == "f64" else np.float32),
                                        64, prec, with_acoustic=True, acoustic_adjoint=field)
            ref[field, prec] = (out[0],) + tuple(a.reshape(db.nspec, 5, 5, 5) for a in out[1:])
    return dict(db=db, src=src, rec=rec, adj=adj, ref=ref)


@pytest.mark.parametrize("exact,assembly,field", [(True, 1, "potential"), (False, 2, "potential"),
                                                  (False, 0, "potential_dot_dot")])
def test_coupled_kernels_match_oracle(problem, exact, assembly, field):
    from specfem3d_amd.adjoint import KernelSolver
    p = problem
    mode = "reference" if field == "potential" else "gradient"
    print(f"[gpu] coupled kernel run exact={exact} assembly={assembly} acoustic_kernels={mode}", flush=True)
    ks = KernelSolver(p["db"], p["src"], p["rec"], NSTEP, nt_dump=64, exact=exact, assembly=assembly,
                      acoustic_kernels=mode)
    seis = ks.forward()
    k = ks.adjoint(p["adj"]) + ks.acoustic_kernels
    ks.close()
    r32, r64 = p["ref"][field, "f32"], p["ref"][field, "f64"]
    if exact and assembly == 1:
        assert np.array_equal(seis, r32[0])
    else:
        assert rel_l2(seis, r32[0]) < 1e-5
    for name, a, b32, b64 in zip(("rho", "mu", "kappa", "rho_ac", "kappa_ac"), k, r32[1:], r64[1:]):
        e32, e64 = rel_l2(a, b32), rel_l2(a, b64)
        print(f"[gpu] {name}_kl rel-L2 vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}", flush=True)
        assert np.abs(a).max() > 0
        assert e64 <= 1e-4


def test_coupled_viscoelastic_kernel_run_matches_oracle(oracle_mod):
    """Fluid over a 3-SLS solid: the snapshots carry the memory variables and the potentials, the
    adjoint couples both ways; forward seismograms bit for bit (exact, ordered), solid and fluid
    kernels within 1e-4 of the FP64 oracle."""
    from specfem3d_amd.adjoint import KernelSolver
    from specfem3d_amd.attenuation import setup_attenuation
    db, src, rec = small_coupled_box(n=4, dt=0.004, nrec=4, seed=3)
    q = np.full((db.nspec, 125), 60.0, np.float32)
    att = setup_attenuation(db, q, qkappa=5 * q, min_resolved_period=0.4, dt=0.004)
    n = 240
    syn = oracle_mod.OracleSolver(db, src, rec, "f32", attenuation=att).run(n)
    adj = (0.3 * syn).astype(np.float32)
    ref = {p: oracle_mod.kernel_run(db, src, rec, n, adj.astype(np.float64 if p == "f64" else np.float32), 48, p,
                                    attenuation=att, with_acoustic=True) for p in ("f32", "f64")}
    print("[gpu] coupled viscoelastic kernel run exact/ordered", flush=True)
    ks = KernelSolver(db, src, rec, n, nt_dump=48, attenuation=att, exact=True, assembly=1)
    seis = ks.forward()
    k = ks.adjoint(adj) + ks.acoustic_kernels
    ks.close()
    assert np.array_equal(seis, ref["f32"][0])
    for name, a, r64 in zip(("rho", "mu", "kappa", "rho_ac", "kappa_ac"), k, ref["f64"][1:]):
        e64 = rel_l2(a, r64.reshape(a.shape))
        print(f"[gpu] coupled viscoelastic {name}_kl rel-L2 vs FP64 oracle {e64:.3e}", flush=True)
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert e64 <= 1e-4
