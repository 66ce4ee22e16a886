"""torch.autograd wrapper of the checkpointed adjoint, for gradient tests only.

``WaveformMisfit.apply(dlnrho, dlnkappa, dlnmu, problem)`` runs the forward solver on the GPU for the
model rho (1 + dlnrho), kappa (1 + dlnkappa), mu (1 + dlnmu) (per local point, (nspec, 5, 5, 5)
float64 tensors), returns the L2 waveform misfit against ``problem.obs`` and, on backward, the
gradient d chi / d dln m = K_m * w J from the checkpointed kernel run (``KernelSolver``). The
gradient is exact at the unperturbed model that the adjoint run linearises about; tests compare its
directional derivative with central finite differences of the misfit.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np
import torch

from . import adjoint as A


@dataclass
class MisfitProblem:
    db: object
    sources: object
    receivers: object
    nstep: int
    obs: np.ndarray                  # observed seismograms (nstep, nrec, 3)
    nt_dump: int = 64
    exact: bool = False
    assembly: int = 0
    tuning: dict = field(default_factory=dict)


def perturbed_database(db, dlnrho=None, dlnkappa=None, dlnmu=None):
    """Copy of ``db`` with rho, kappa, mu scaled by (1 + dln m) at every local point. The mass
    matrix is rebuilt when rho changes; Stacey impedances are left as they are."""
    d = copy.copy(db)
    for name, dl in (("rhostore", dlnrho), ("kappastore", dlnkappa), ("mustore", dlnmu)):
        if dl is None:
            continue
        a = getattr(db, name)
        f = 1.0 + np.asarray(dl, np.float64).reshape(a.shape)
        setattr(d, name, (a.astype(np.float64) * f).astype(np.float32))
    if dlnrho is not None:
        d.finalize(db.dt, stacey=db.stacey)
    return d


class WaveformMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dlnrho, dlnkappa, dlnmu, problem: MisfitProblem):
        p = problem
        conv = (lambda t: t.detach().cpu().numpy())
        db = perturbed_database(p.db, conv(dlnrho), conv(dlnkappa), conv(dlnmu))
        ks = A.KernelSolver(db, p.sources, p.receivers, p.nstep, nt_dump=p.nt_dump, exact=p.exact,
                            assembly=p.assembly, tuning=p.tuning)
        try:
            syn = ks.forward()
            chi = A.waveform_misfit(syn, p.obs, db.dt)
            rk, mk, kk = ks.adjoint(A.waveform_adjoint_source(syn, p.obs))
        finally:
            ks.close()
        K = A.finalize_kernels(db, rk, mk, kk)
        wJ = A.quadrature_weights(db)
        grads = [torch.from_numpy((K[k].astype(np.float64) * wJ)) for k in ("rho", "kappa", "mu")]
        ctx.save_for_backward(*grads)
        return torch.tensor(chi, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gr, gk, gm = ctx.saved_tensors
        return g * gr, g * gk, g * gm, None
