"""Checkpointed adjoint (sensitivity-kernel) runs on the GPU, and the host-side pieces around them.

The reference computes kernels with SIMULATION_TYPE=3 either by reconstructing the forward field
backward in time from its last frame plus stored boundary records (iterate_time.F90:233-290), or,
with UNDO_ATTENUATION_AND_OR_PML, by forward checkpointing and recomputation
(iterate_time_undoatt.F90:29-632). ``KernelSolver`` follows the second scheme through the C-ABI:

* ``forward()``: the SIMULATION_TYPE=1 + SAVE_FORWARD pass. The forward field is stepped from rest
  and its state is saved at the start of every subset of ``nt_dump`` steps
  (save_forward_arrays_undoatt) -- here into HBM (``sf_snapshot_save``) instead of files;
* ``adjoint(adj)``: subsets in reverse; each restores a snapshot (``sf_snapshot_load`` +
  ``compute_strain_cuda``), recomputes the subset keeping every step's displacement in an HBM
  history (``sf_history_store``), then steps the adjoint field through the subset with the
  reconstructed displacement read back time-reversed (``sf_history_load``) and accumulates the
  kernels (``compute_kernels_elastic_cuda``, and ``compute_kernels_acoustic_cuda`` in fluid
  elements) after each adjoint step.

``finalize_kernels`` applies save_adjoint_kernels.f90:276-412 (rho, mu, kappa, rhop, alpha, beta),
``finalize_acoustic_kernels`` save_adjoint_kernels.f90:188-222 (rho_ac, kappa_ac, rhop_ac,
alpha_ac); ``write_kernel_bin`` writes them as Fortran unformatted records. The torch.autograd wrapper used by
the gradient tests lives in ``specfem3d_amd.autograd``.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import abi
from .gll import get_gll
from .solver import GPUSolver


# ------------------------------------------------------------------------------ host-side pieces
def waveform_misfit(syn: np.ndarray, obs: np.ndarray, dt: float) -> float:
    """L2 waveform misfit chi = 1/2 sum_r sum_t |syn - obs|^2 dt (FP64 accumulation)."""
    r = syn.astype(np.float64) - obs.astype(np.float64)
    return 0.5 * float((r * r).sum()) * float(dt)


def waveform_adjoint_source(syn: np.ndarray, obs: np.ndarray) -> np.ndarray:
    """Adjoint source of the L2 waveform misfit in forward time, (nstep, nrec, 3): syn - obs.

    This is what the ``SEM/NET.STA.BX?.adj`` files hold; the solver reads it time-reversed
    (compute_arrays_source.f90:203-291)."""
    return (syn.astype(np.float64) - obs.astype(np.float64)).astype(np.float32)


def quadrature_weights(db) -> np.ndarray:
    """w_i w_j w_k J at every local point, (nspec, 5, 5, 5) [e, k, j, i]: the volume weights that turn
    a kernel into a directional derivative, d chi = sum K * w J * dln m."""
    g = get_gll()
    w = np.asarray(g.wgll, np.float64)
    w3 = w[None, None, :] * w[None, :, None] * w[:, None, None]
    out = np.empty((db.nspec, 5, 5, 5), np.float64)
    irr = db.irregular_element_number
    reg = irr == 0
    out[reg] = w3[None] * float(db.jacobian_regular)
    if (~reg).any():
        jac = np.asarray(db.jacobian, np.float64).reshape(-1, 5, 5, 5)
        out[~reg] = w3[None] * jac[irr[~reg] - 1]
    return out


def finalize_kernels(db, rho_kl, mu_kl, kappa_kl, mustore=None, kappastore=None) -> dict:
    """save_adjoint_kernels.f90:276-412 (isotropic, elastic elements): rho = -rho K_rho,
    mu = -2 mu K_mu, kappa = -kappa K_kappa, rhop = rho + kappa + mu, beta and alpha kernels.
    rho is recovered as rho_vs^2 / mu as the reference does."""
    shp = (db.nspec, 5, 5, 5)
    mu = np.asarray(db.mustore if mustore is None else mustore, np.float32).reshape(shp)
    kap = np.asarray(db.kappastore if kappastore is None else kappastore, np.float32).reshape(shp)
    rvs = np.asarray(db.rho_vs, np.float32).reshape(shp)
    rk, mk, kk = (np.asarray(a, np.float32).reshape(shp) for a in (rho_kl, mu_kl, kappa_kl))
    el = db.ispec_is_elastic
    out = {k: np.zeros(shp, np.float32) for k in ("rho", "mu", "kappa", "rhop", "alpha", "beta")}
    with np.errstate(divide="ignore", invalid="ignore"):
        rhol = rvs * rvs / mu
        rho = -rhol * rk
        mu_k = np.float32(-2.0) * mu * mk
        kap_k = -kap * kk
        out["rho"][el] = rho[el]
        out["mu"][el] = mu_k[el]
        out["kappa"][el] = kap_k[el]
        out["rhop"][el] = (rho + kap_k + mu_k)[el]
        out["beta"][el] = (np.float32(2.0) * (mu_k - np.float32(4.0 / 3.0) * mu / kap * kap_k))[el]
        out["alpha"][el] = (np.float32(2.0) * (np.float32(1.0) + np.float32(4.0 / 3.0) * mu / kap) * kap_k)[el]
    return out


def finalize_acoustic_kernels(db, rho_ac_kl, kappa_ac_kl) -> dict:
    """save_adjoint_kernels.f90:188-222 (fluid elements): the accumulated integrals are kept positive
    and become the relative kernels rho_ac = rho_ac_kl / rho, kappa_ac = kappa_ac_kl / kappa, with
    rhop_ac = rho_ac + kappa_ac and alpha_ac = 2 kappa_ac."""
    shp = (db.nspec, 5, 5, 5)
    rho = np.asarray(db.rhostore, np.float32).reshape(shp)
    kap = np.asarray(db.kappastore, np.float32).reshape(shp)
    rk, kk = (np.asarray(a, np.float32).reshape(shp) for a in (rho_ac_kl, kappa_ac_kl))
    ac = db.ispec_is_acoustic
    out = {k: np.zeros(shp, np.float32) for k in ("rho_ac", "kappa_ac", "rhop_ac", "alpha_ac")}
    r = (np.float32(1.0) / rho[ac]) * rk[ac]
    k = (np.float32(1.0) / kap[ac]) * kk[ac]
    out["rho_ac"][ac] = r
    out["kappa_ac"][ac] = k
    out["rhop_ac"][ac] = r + k
    out["alpha_ac"][ac] = np.float32(2.0) * k
    return out


def write_kernel_bin(path: str, arr: np.ndarray) -> None:
    """One Fortran unformatted sequential record (4-byte length markers), as save_adjoint_kernels
    writes ``proc000000_<name>_kernel.bin`` (the array in (i,j,k,ispec) order)."""
    a = np.ascontiguousarray(arr, np.float32).reshape(-1)
    n = np.array([a.nbytes], np.int32)
    with open(path, "wb") as f:
        f.write(n.tobytes())
        f.write(a.tobytes())
        f.write(n.tobytes())


def read_kernel_bin(path: str) -> np.ndarray:
    raw = np.fromfile(path, np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    if n + 8 != raw.size or int(raw[-4:].view(np.int32)[0]) != n:
        raise ValueError(f"{path}: not a single Fortran unformatted record")
    return raw[4:4 + n].view(np.float32).copy()


# ------------------------------------------------------------------------------ GPU driver
class KernelSolver(GPUSolver):
    """Checkpointed sensitivity-kernel run on the GPU (SIMULATION_TYPE=3, undo-attenuation loop).

    Elastic, viscoelastic and coupled acoustic-elastic meshes (sources and receivers in solid or
    fluid elements); the forward/adjoint field is ``FORWARD_OR_ADJOINT`` 1 and the reconstructed forward
    field 3, as in the reference. Each step runs the acoustic domain first, then the elastic one
    (compute_forces_acoustic_calling, then compute_forces_viscoelastic_calling).

    ``acoustic_kernels="reference"`` accumulates the fluid kernels as the reference does, from the
    adjoint potential (compute_kernels.f90:269,325). Because the adjoint potential is driven by the
    adjoint displacement through the coupling, the misfit gradient needs its second time derivative
    instead: ``"gradient"`` uses the adjoint potential_dot_dot (same arithmetic otherwise)."""

    def __init__(self, db, sources, receivers, nstep: int, nt_dump: int = 64, attenuation=None,
                 exact: bool = False, assembly: int = 0, device: int = 0, tuning: dict | None = None,
                 acoustic_kernels: str = "reference"):
        if acoustic_kernels not in ("reference", "gradient"):
            raise ValueError("acoustic_kernels must be 'reference' or 'gradient'")
        tuning = dict(tuning or {})
        tuning["acoustic_kernel_field"] = int(acoustic_kernels == "gradient")
        if not db.nspec_elastic:
            raise NotImplementedError("kernel runs need solid elements (sources and receivers are elastic)")
        if receivers is None or receivers.n == 0:
            raise ValueError("a kernel run needs receivers (the adjoint sources)")
        super().__init__(db, sources, receivers, nstep=nstep, attenuation=attenuation, exact=exact,
                         assembly=assembly, simulation_type=3, nlength=nstep, device=device, tuning=tuning,
                         undo_attenuation=True)
        self.nt_dump = max(1, min(int(nt_dump), self.nstep))
        self.nsub = math.ceil(self.nstep / self.nt_dump)
        self.forward_done = False
        self.acoustic_kernels = None   # raw (rho_ac_kl, kappa_ac_kl) of the last adjoint() in fluid meshes

    def _subset_len(self, slot: int) -> int:
        return self.nt_dump if slot < self.nsub - 1 else self.nstep - (self.nsub - 1) * self.nt_dump

    def _acoustic_step(self, foa: int, fwd_it: int = 0, it: int = 0, adj: np.ndarray | None = None):
        """Fluid part of one step of field ``foa`` (compute_forces_acoustic_calling /
        _backward_calling, undo-attenuation Stacey): free surface, stiffness, Stacey, the
        solid-to-fluid coupling and the sources in fluid elements (the CMT sources at forward step
        ``fwd_it`` on the reconstructed field, the adjoint sources of step ``it`` on the adjoint field),
        corrector, free surface."""
        L, MP = self.L, self.MP
        f, zero = self._ref(foa), self._ref(0)
        dt2 = C.byref(self._c["dt2"])
        ncf = C.c_int(self.db.coupling_faces.n)
        L.acoustic_enforce_free_surf_cuda_(MP, zero, f)
        for iphase in (1, 2):
            ip = C.c_int(iphase)
            L.compute_forces_acoustic_cuda_(MP, C.byref(ip), zero, zero, f)
            if iphase == 1:
                L.compute_stacey_acoustic_undoatt_cuda_(MP, C.byref(ip), f)
                L.compute_coupling_ac_el_cuda_(MP, C.byref(ip), C.byref(ncf), f)
                if foa == 3 and self.stf_table is not None:
                    stf = np.ascontiguousarray(self.stf_table[fwd_it - 1], np.float64)
                    L.compute_add_sources_ac_s3_cuda_(MP, C.byref(C.c_int(len(stf))),
                                                      stf.ctypes.data_as(C.c_void_p), zero)
                if foa == 1 and adj is not None and it < self.nstep:
                    nrec, ntr = C.c_int(self.receivers.n), C.c_int(self.nstep)
                    L.add_sources_ac_sim_2_or_3_cuda_(MP, abi.ptr(adj), C.byref(nrec), C.byref(nrec), C.byref(ntr),
                                                      C.byref(C.c_int(it)))
        L.kernel_3_acoustic_cuda_(MP, dt2, dt2, f)
        L.acoustic_enforce_free_surf_cuda_(MP, zero, f)

    def _b_step(self, fwd_it: int):
        """One step of the reconstructed field (compute_forces_viscoelastic_backward_calling with
        b_deltat = +DT, prepare_timerun.F90:418-422), forward-time step index ``fwd_it``."""
        L, MP = self.L, self.MP
        three, zero = self._ref(3), self._ref(0)
        dt, dt2, dtsq2 = (C.byref(self._c[k]) for k in ("dt", "dt2", "dtsq2"))
        if self.acoustic:
            L.update_displacement_ac_cuda_(MP, dt, dt2, dtsq2, three)
        L.update_displacement_cuda_(MP, dt, dt2, dtsq2, three)
        if self.acoustic:
            self._acoustic_step(3, fwd_it=fwd_it)
        strain = C.c_int(1)
        att = C.c_int(int(self.att is not None))
        ncf = C.c_int(self.db.coupling_faces.n)
        for iphase in (1, 2):
            ip = C.c_int(iphase)
            L.compute_forces_viscoelastic_cuda_(MP, C.byref(ip), dt, zero, zero, C.byref(strain), C.byref(att), three)
            if iphase == 1:
                L.compute_stacey_viscoelastic_undoatt_cuda_(MP, C.byref(ip), three)
                if self.acoustic:
                    L.compute_coupling_el_ac_cuda_(MP, C.byref(ip), C.byref(ncf), three)
                if self.stf_table is not None:
                    stf = np.ascontiguousarray(self.stf_table[fwd_it - 1], np.float64)
                    L.compute_add_sources_el_s3_cuda_(MP, stf.ctypes.data_as(C.c_void_p), C.byref(C.c_int(len(stf))))
        L.kernel_3_a_cuda_(MP, dt2, dt2, zero, three)

    def _adj_step(self, it: int, adj: np.ndarray):
        """One step of the adjoint field (compute_forces_viscoelastic_calling at SIMULATION_TYPE=3)."""
        L, MP = self.L, self.MP
        one, zero = self._ref(1), self._ref(0)
        dt, dt2, dtsq2 = (C.byref(self._c[k]) for k in ("dt", "dt2", "dtsq2"))
        if self.acoustic:
            L.update_displacement_ac_cuda_(MP, dt, dt2, dtsq2, one)
        L.update_displacement_cuda_(MP, dt, dt2, dtsq2, one)
        if self.acoustic:
            self._acoustic_step(1, it=it, adj=adj)
        strain = C.c_int(1)
        att = C.c_int(int(self.att is not None))
        nrec = C.c_int(self.receivers.n)
        ntr = C.c_int(self.nstep)
        ncf = C.c_int(self.db.coupling_faces.n)
        for iphase in (1, 2):
            ip = C.c_int(iphase)
            L.compute_forces_viscoelastic_cuda_(MP, C.byref(ip), dt, zero, zero, C.byref(strain), C.byref(att), one)
            if iphase == 1:
                L.compute_stacey_viscoelastic_undoatt_cuda_(MP, C.byref(ip), one)
                if self.acoustic:
                    L.compute_coupling_el_ac_cuda_(MP, C.byref(ip), C.byref(ncf), one)
                if it < self.nstep:
                    L.add_sources_el_sim_type_2_or_3_(MP, abi.ptr(adj), C.byref(nrec), C.byref(nrec), C.byref(ntr),
                                                      C.byref(C.c_int(it)))
        L.kernel_3_a_cuda_(MP, dt2, dt2, zero, one)

    def forward(self) -> np.ndarray:
        """Forward pass with a snapshot per subset; returns the forward seismograms (nstep, nrec, 3)."""
        L, MP = self.L, self.MP
        self._zero_field(3)
        self.seismo_current = 0
        it = 0
        for slot in range(self.nsub):
            L.sf_snapshot_save(MP, C.c_int(3), C.c_int(slot))
            for _ in range(self._subset_len(slot)):
                it += 1
                self._b_step(it)
                self._record(it)
        self.forward_done = True
        return self.seis_d[: self.seismo_current].copy()

    def adjoint(self, adj: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Adjoint pass driven by ``adj`` (nstep, nrec, 3) in forward time; returns the raw kernel
        time integrals (rho_kl, mu_kl, kappa_kl) over (nspec, 5, 5, 5)."""
        if not self.forward_done:
            raise RuntimeError("KernelSolver.adjoint() needs forward() first (no snapshots saved)")
        adj = np.ascontiguousarray(adj, np.float32)
        if adj.shape != (self.nstep, self.receivers.n, 3):
            raise ValueError(f"adjoint source must be (nstep, nrec, 3) = {(self.nstep, self.receivers.n, 3)}")
        L, MP = self.L, self.MP
        self._keep_adj = adj
        self._zero_field(1)
        self._zero_kernels()
        L.sf_history_reserve(MP, C.c_int(self.nt_dump))
        one, three = self._ref(1), self._ref(3)
        it = 0
        for i in range(1, self.nsub + 1):
            slot = self.nsub - i
            n_i = self._subset_len(slot)
            L.sf_snapshot_load(MP, C.c_int(3), C.c_int(slot))
            L.compute_strain_cuda_(MP, three)
            base = slot * self.nt_dump
            for m in range(n_i):
                self._b_step(base + m + 1)
                L.sf_history_store(MP, C.c_int(3), C.c_int(m))
            L.compute_strain_cuda_(MP, one)
            dtk = C.c_float(self.deltat)
            for m in range(n_i):
                it += 1
                L.sf_history_load(MP, C.c_int(3), C.c_int(n_i - 1 - m))
                self._adj_step(it, adj)
                L.compute_kernels_elastic_cuda_(MP, C.byref(dtk))
                if self.acoustic:
                    L.compute_kernels_acoustic_cuda_(MP, C.byref(dtk))
        L.sf_history_reserve(MP, C.c_int(0))
        nloc = self.db.nspec * 125
        rk, mk, kk = (np.zeros(nloc, np.float32) for _ in range(3))
        L.transfer_kernels_el_to_host_(MP, abi.ptr(rk), abi.ptr(mk), abi.ptr(kk), None, C.byref(C.c_int(self.db.nspec)))
        shp = (self.db.nspec, 5, 5, 5)
        if self.acoustic:
            ra, ka = np.zeros(nloc, np.float32), np.zeros(nloc, np.float32)
            L.transfer_kernels_ac_to_host_(MP, abi.ptr(ra), abi.ptr(ka), C.byref(C.c_int(self.db.nspec)))
            self.acoustic_kernels = (ra.reshape(shp), ka.reshape(shp))
        return rk.reshape(shp), mk.reshape(shp), kk.reshape(shp)

    # -- helpers
    def _record(self, it: int):
        zero = self._ref(0)
        self.seismo_current += 1
        self.L.compute_seismograms_cuda_(
            self.MP, abi.ptr(self.seis_d), abi.ptr(self.seis_v), abi.ptr(self.seis_a), abi.ptr(self.seis_p),
            C.byref(C.c_int(self.seismo_current)), C.byref(C.c_int(self.nlength)), C.byref(C.c_int(it)),
            C.byref(C.c_int(self.nstep)), C.byref(C.c_int(0)), C.byref(C.c_int(1)), zero)

    def _zero_field(self, foa: int):
        n3 = 3 * self.db.nglob
        z = np.zeros(n3, np.float32)
        A = abi.Args()
        if foa == 3:
            self.L.transfer_b_fields_to_device_(A.i(n3), abi.ptr(z), abi.ptr(z), abi.ptr(z), self.MP)
        else:
            self.L.transfer_fields_el_to_device_(A.i(n3), abi.ptr(z), abi.ptr(z), abi.ptr(z), self.MP)
        if self.acoustic:
            n = A.i(self.db.nglob)
            xfer = self.L.transfer_b_fields_ac_to_device_ if foa == 3 else self.L.transfer_fields_ac_to_device_
            xfer(n, abi.ptr(z), abi.ptr(z), abi.ptr(z), self.MP)

    def _zero_kernels(self):
        self.L.sf_zero_kernels(self.MP)


# ------------------------------------------------------------------------------ time reversal
class ReversalKernelSolver(GPUSolver):
    """SIMULATION_TYPE=3 by time-reversed reconstruction, the reference's classic kernel run
    (iterate_time.F90:233-290, compute_forces_viscoelastic_calling_routine.F90:694-870).

    The forward run (``record_forward``: SIMULATION_TYPE=1 with SAVE_FORWARD) keeps the Stacey
    traction record of every step (write_abs of compute_stacey_viscoelastic.F90:466-478) and its
    last frame (save_forward_arrays). The kernel run starts the backward field from that frame and
    steps it with b_deltat = -DT (prepare_timerun.F90:424-429), replaying the records in reverse
    (read_abs NSTEP-it+1) and adding the sources at time (NSTEP-it) DT - t0
    (compute_add_sources_viscoelastic.F90:650-690), while the adjoint field steps forward. The
    checkpointed ``KernelSolver`` replaces this scheme; it stays available for drop-in hosts and is
    tested by reconstructing the forward wavefield."""

    def __init__(self, db, sources, receivers, nstep: int, exact: bool = False, assembly: int = 0, device: int = 0,
                 tuning: dict | None = None):
        if db.nspec_acoustic:
            raise NotImplementedError("time-reversed reconstruction is driven for elastic domains only")
        super().__init__(db, sources, receivers, nstep=nstep, exact=exact, assembly=assembly, simulation_type=3,
                         nlength=nstep, device=device, tuning=tuning)
        self.b_c = {k: C.c_float(v) for k, v in dict(
            dt=np.float32(-self.deltat), dt2=np.float32(np.float32(-self.deltat) / np.float32(2)),
            dtsq2=np.float32(np.float32(-self.deltat) * np.float32(-self.deltat) / np.float32(2))).items()}

    @staticmethod
    def record_forward(db, sources, receivers, nstep: int, exact: bool = False, assembly: int = 0,
                       keep_displ: bool = False, device: int = 0):
        """Forward run with SAVE_FORWARD: returns (seismograms, Stacey records (nstep, 3*25*nfaces),
        last frame (displ, veloc, accel)[, displacement of every step])."""
        g = GPUSolver(db, sources, receivers, nstep=nstep, exact=exact, assembly=assembly, save_forward=True,
                      device=device)
        recs, disp = [], []
        for _ in range(nstep):
            g.step()
            recs.append(g.b_absorb_field.copy())
            if keep_displ:
                disp.append(g.fields()[0])
        seis = g.seis_d[: g.seismo_current].copy()
        last = g.fields()
        g.close()
        return (seis, np.stack(recs), last) + ((np.stack(disp),) if keep_displ else ())

    def start_from(self, last):
        """read_forward_arrays: the backward field starts at the forward run's last frame."""
        n3 = 3 * self.db.nglob
        self._last = [np.ascontiguousarray(a, np.float32) for a in last]
        self.L.transfer_b_fields_to_device_(C.byref(C.c_int(n3)), *(abi.ptr(a) for a in self._last), self.MP)
        self.it = 0

    def backward_step(self, records: np.ndarray, exact_reversal: bool = False):
        """One reversed step of the backward field (FORWARD_OR_ADJOINT = 3).

        As the reference indexes them, step ``it`` replays the record of forward step NSTEP-it+1
        (read_abs) and adds the sources at (NSTEP-it) DT - t0. Undoing forward step NSTEP-it of the
        Newmark scheme exactly needs that step's own record and source time, one step earlier:
        ``exact_reversal`` uses those (the reconstruction is then exact to rounding)."""
        L, MP = self.L, self.MP
        self.it += 1
        it = self.it
        k = self.nstep - it - (1 if exact_reversal else 0)    # 0-based forward step of the record/STF
        three, zero = self._ref(3), self._ref(0)
        bdt, bdt2, bdtsq2 = (C.byref(self.b_c[k]) for k in ("dt", "dt2", "dtsq2"))
        L.update_displacement_cuda_(MP, bdt, bdt2, bdtsq2, three)
        strain, att = C.c_int(1), C.c_int(0)
        for iphase in (1, 2):
            ip = C.c_int(iphase)
            L.compute_forces_viscoelastic_cuda_(MP, C.byref(ip), bdt, zero, zero, C.byref(strain), C.byref(att), three)
            if iphase == 1:
                self._rec = np.ascontiguousarray(records[k], np.float32)
                L.compute_stacey_viscoelastic_cuda_(MP, C.byref(ip), abi.ptr(self._rec), three)
                if self.stf_table is not None:
                    stf = np.ascontiguousarray(self.stf_table[k], np.float64)
                    L.compute_add_sources_el_s3_cuda_(MP, stf.ctypes.data_as(C.c_void_p), C.byref(C.c_int(len(stf))))
        L.kernel_3_a_cuda_(MP, bdt2, bdt2, zero, three)

    def b_displ(self) -> np.ndarray:
        n3 = 3 * self.db.nglob
        out = np.zeros(n3, np.float32)
        self.L.transfer_b_displ_from_device_(C.byref(C.c_int(n3)), abi.ptr(out), self.MP)
        return out
