"""GPU checkpointed kernel run (KernelSolver through the C-ABI) against the oracle, and the autograd
wrapper's gradient against central finite differences of the misfit computed on the GPU.

Targets: sensitivity kernels within 1e-4 aggregate rel-L2 of the FP64 oracle; the reconstructed
forward seismograms equal the FP32 oracle bit for bit in exact/ordered mode.
"""
import copy

import numpy as np
import pytest

from conftest import rel_l2, small_elastic_box

pytestmark = pytest.mark.gpu

NSTEP = 400


@pytest.fixture(scope="module")
def problem(oracle_mod):
    from specfem3d_amd import adjoint as A
    db, src, rec = small_elastic_box(n=4, dt=0.004, nrec=6)
    L = 2000.0
    ib = db.ibool.reshape(db.nspec, 125)
    xl, yl, zl = db.xstore[ib], db.ystore[ib], db.zstore[ib]
    s = L / 6
    phi = np.exp(-((xl - L / 2) ** 2 + (yl - L / 2) ** 2 + (zl + L / 2) ** 2) / (2 * s * s)).reshape(db.nspec, 5, 5, 5)
    d = copy.copy(db)
    d.mustore = (db.mustore.astype(np.float64) * (1 + 0.03 * phi.reshape(db.mustore.shape))).astype(np.float32)
    obs = oracle_mod.OracleSolver(d, src, rec, "f32").run(NSTEP)
    syn = oracle_mod.OracleSolver(db, src, rec, "f32").run(NSTEP)
    adj = A.waveform_adjoint_source(syn, obs)
    ref = {}
    for prec in ("f32", "f64"):
        seis, rk, mk, kk = oracle_mod.kernel_run(db, src, rec, NSTEP, adj.astype(np.float64 if prec == "f64" else np.float32),
                                                 64, prec)
        ref[prec] = (seis, rk.reshape(db.nspec, 5, 5, 5), mk.reshape(db.nspec, 5, 5, 5), kk.reshape(db.nspec, 5, 5, 5))
    return dict(db=db, src=src, rec=rec, phi=phi, obs=obs, syn=syn, adj=adj, ref=ref)


def run_kernels(p, exact, assembly, nt_dump=64):
    from specfem3d_amd.adjoint import KernelSolver
    print(f"[gpu] kernel run exact={exact} assembly={assembly} nt_dump={nt_dump}", flush=True)
    ks = KernelSolver(p["db"], p["src"], p["rec"], NSTEP, nt_dump=nt_dump, exact=exact, assembly=assembly)
    seis = ks.forward()
    k = ks.adjoint(p["adj"])
    ks.close()
    return seis, k


@pytest.mark.parametrize("exact,assembly", [(True, 1), (False, 0), (False, 2)])
def test_kernels_match_oracle(problem, exact, assembly):
    p = problem
    seis, k = run_kernels(p, exact, assembly)
    if exact and assembly == 1:
        assert np.array_equal(seis, p["ref"]["f32"][0])
    else:
        assert rel_l2(seis, p["ref"]["f32"][0]) < 1e-5
    for name, a, r32, r64 in zip(("rho", "mu", "kappa"), k, p["ref"]["f32"][1:], p["ref"]["f64"][1:]):
        e32, e64 = rel_l2(a, r32), rel_l2(a, r64)
        print(f"[gpu] {name}_kl rel-L2 vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}", flush=True)
        assert np.abs(a).max() > 0
        assert e64 <= 1e-4


def test_checkpoint_interval_invariance_on_gpu(problem):
    p = problem
    _, k1 = run_kernels(p, True, 1, nt_dump=64)
    _, k2 = run_kernels(p, True, 1, nt_dump=NSTEP)
    for a, b in zip(k1, k2):
        assert np.array_equal(a, b)


def test_autograd_gradient_matches_central_fd(problem):
    import torch
    from specfem3d_amd import adjoint as A
    from specfem3d_amd.autograd import MisfitProblem, WaveformMisfit, perturbed_database
    from specfem3d_amd.solver import GPUSolver
    p = problem
    db = p["db"]
    prob = MisfitProblem(db=db, sources=p["src"], receivers=p["rec"], nstep=NSTEP, obs=p["obs"])
    shp = (db.nspec, 5, 5, 5)
    z = [torch.zeros(shp, dtype=torch.float64, requires_grad=True) for _ in range(3)]
    chi = WaveformMisfit.apply(z[0], z[1], z[2], prob)
    chi.backward()
    phi = torch.from_numpy(p["phi"])
    h = 1e-3
    for i, name in enumerate(("rho", "kappa", "mu")):
        dk = float((z[i].grad * phi).sum())

        def misfit(sign):
            dl = [None, None, None]
            dl[i] = sign * h * p["phi"]
            d = perturbed_database(db, *dl)
            g = GPUSolver(d, p["src"], p["rec"], nstep=NSTEP, exact=False, assembly=2)
            s = g.run().copy()
            g.close()
            return A.waveform_misfit(s, p["obs"], db.dt)

        fd = (misfit(+1) - misfit(-1)) / (2 * h)
        print(f"[gpu] d chi / d ln {name} along phi: autograd {dk:.6e}, central FD {fd:.6e}", flush=True)
        assert abs(dk - fd) <= 5e-3 * abs(fd)


def test_viscoelastic_kernel_run_matches_oracle(oracle_mod):
    """3-SLS attenuation in both fields: the snapshots carry the memory variables, the strain is
    recomputed from the restored displacement (with its trace plane), as in the reference."""
    from specfem3d_amd.adjoint import KernelSolver
    from specfem3d_amd.attenuation import setup_attenuation
    db, src, rec = small_elastic_box(n=3, dt=0.004, nrec=4, seed=2)
    q = np.full((db.nspec, 125), 60.0, np.float32)
    att = setup_attenuation(db, q, qkappa=5 * q, min_resolved_period=0.4, dt=0.004)
    n = 240
    syn = oracle_mod.OracleSolver(db, src, rec, "f32", attenuation=att).run(n)
    adj = (0.3 * syn).astype(np.float32)
    ref = {p: oracle_mod.kernel_run(db, src, rec, n, adj.astype(np.float64 if p == "f64" else np.float32), 48, p,
                                    attenuation=att) for p in ("f32", "f64")}
    print("[gpu] viscoelastic kernel run exact/ordered", flush=True)
    ks = KernelSolver(db, src, rec, n, nt_dump=48, attenuation=att, exact=True, assembly=1)
    seis = ks.forward()
    k = ks.adjoint(adj)
    ks.close()
    assert np.array_equal(seis, ref["f32"][0])
    for name, a, r32, r64 in zip(("rho", "mu", "kappa"), k, ref["f32"][1:], ref["f64"][1:]):
        e32 = rel_l2(a, r32.reshape(a.shape))
        e64 = rel_l2(a, r64.reshape(a.shape))
        print(f"[gpu] viscoelastic {name}_kl rel-L2 vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e}", flush=True)
        assert np.abs(a).max() > 0
        assert e64 <= 1e-4
