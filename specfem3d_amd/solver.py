"""Host driver of the GPU solver loop through the Fortran-callable C-ABI.

``GPUSolver`` replays, call for call, what the reference's Fortran host does at NPROC=1 in
GPU_MODE:

* set-up: ``prepare_GPU`` (``src/specfem3D/prepare_gpu.f90:29-352``): initialize_gpu_device,
  prepare_constants_device, prepare_fields_acoustic_device, prepare_fields_elastic_device,
  (kernel runs: prepare_fields_elastic_adj_dev, prepare_sim2_or_3_const_device),
  transfer_fields_*_to_device;
* each time step: ``iterate_time.F90:219-330`` -- update_displ_Newmark (acoustic, then
  elastic), ``compute_forces_acoustic_forward_calling``
  (``compute_forces_acoustic_calling_routine.f90:54-343``), ``compute_forces_viscoelastic_calling``
  (``compute_forces_viscoelastic_calling_routine.F90:30-435``), write_seismograms;
* end: the last seismogram chunk comes back from ``compute_seismograms_cuda``.

Arrays are handed over in the reference's Fortran layouts (column-major, 1-based indices).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .gll import get_gll
from .mesh.database import Database


def morton_order(db, elems: np.ndarray) -> np.ndarray:
    """``elems`` sorted along a Morton (Z-order) curve of their centres: elements that share points
    end up close together in the launch order (their shared rows meet in L2)."""
    elems = np.asarray(elems, np.int64)
    if elems.size == 0:
        return elems.astype(np.int32)
    ib = db.ibool.reshape(db.nspec, 125)[elems]
    c = np.stack([db.xstore[ib].mean(1), db.ystore[ib].mean(1), db.zstore[ib].mean(1)], 1).astype(np.float64)
    lo, hi = c.min(0), c.max(0)
    q = np.floor((c - lo) / np.maximum(hi - lo, 1e-30) * 1023.0 + 0.5).astype(np.uint64)
    key = np.zeros(len(elems), np.uint64)
    for bit in range(10):
        for d in range(3):
            key |= ((q[:, d] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + d)
    return elems[np.argsort(key, kind="stable")].astype(np.int32)


def brick_order(db, elems: np.ndarray, by: int = 2, bz: int = 2) -> np.ndarray:
    """``elems`` in slab order: (z // bz, y // by, x, y % by, z % bz) of their centres on the element
    grid. Consecutive launches then walk by * bz rows of the natural numbering side by side, so the
    per-element state arrays are still read as a few sequential streams, while the neighbours across
    y and z faces run within a few launches of each other (their shared accumulator lines meet in
    cache)."""
    elems = np.asarray(elems, np.int64)
    if elems.size == 0:
        return elems.astype(np.int32)
    ib = db.ibool.reshape(db.nspec, 125)[elems]
    c = np.stack([db.xstore[ib].mean(1), db.ystore[ib].mean(1), db.zstore[ib].mean(1)], 1).astype(np.float64)
    # element grid indices from the centres and the smallest element extent along each axis
    ext = np.stack([db.xstore[ib].max(1) - db.xstore[ib].min(1), db.ystore[ib].max(1) - db.ystore[ib].min(1),
                    db.zstore[ib].max(1) - db.zstore[ib].min(1)], 1).astype(np.float64)
    h = np.maximum(ext.min(0), 1e-30)
    ix, iy, iz = (np.floor((c[:, d] - c[:, d].min()) / h[d] + 0.5).astype(np.int64) for d in range(3))
    key = np.lexsort((iz % bz, iy % by, ix, iy // by, iz // bz))
    return elems[key].astype(np.int32)


def _fortran_phase_list(elems: np.ndarray) -> tuple[int, np.ndarray]:
    """phase_ispec_inner_*(num, 2): column 1 outer (none at NPROC=1), column 2 inner, 1-based."""
    n = max(len(elems), 1)
    a = np.zeros((2, n), np.int32)
    a[1, :len(elems)] = elems + 1
    return n, a.ravel()


class GPUSolver:
    """Forward solver on the GPU (SIMULATION_TYPE = 1), optionally saving boundary records."""

    def __init__(self, db: Database, sources=None, receivers=None, nstep: int = 1000,
                 attenuation=None, exact: bool = True, assembly: int = 0, simulation_type: int = 1,
                 save_forward: bool = False, save_seismograms=("d",), nlength: int | None = None,
                 device: int = 0, tuning: dict | None = None, undo_attenuation: bool = False,
                 element_order: str | None = None, pressure_trick: bool = False):
        self.L = abi.load()
        self.db = db
        self.sources = sources
        self.receivers = receivers
        self.nstep = int(nstep)
        self.sim = simulation_type
        self.att = attenuation
        self.mp = C.c_long(0)
        self.MP = C.byref(self.mp)
        self.elastic = db.nspec_elastic > 0
        self.acoustic = db.nspec_acoustic > 0
        self.it = 0
        self.seismo_current = 0
        self.nlength = int(nlength or nstep)
        self.save = save_seismograms
        self.tuning = dict(tuning or {})
        self.undo_att = bool(undo_attenuation)
        self.element_order = element_order
        # USE_TRICK_FOR_BETTER_PRESSURE: pressure recorded as -potential (the fluid sources then use the
        # second derivative of their STF, see sources.setup_cmt_sources)
        self.pressure_trick = bool(pressure_trick)
        self._prepare(exact, assembly, save_forward, device)

    # ------------------------------------------------------------------------------ set-up
    def _prepare(self, exact, assembly, save_forward, device):
        import os
        import sys
        import time
        t0 = time.perf_counter()
        verbose = os.environ.get("SF_VERBOSE_PREPARE") == "1"

        def stage(name):
            if verbose:
                print(f"[prepare] {name} done at {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        L, db, A = self.L, self.db, abi.Args()
        self._keep = A
        ndev = C.c_int(0)
        L.initialize_gpu_device_(A.i(device), C.byref(ndev))
        g = get_gll()
        nirr = db.nspec_irregular
        jac = [db.jac9[:, t, :] for t in range(9)] if nirr else [np.zeros(1, np.float32)] * 9
        # absorbing faces: Stacey faces, or with C-PML the outer faces of the layers (Dirichlet)
        f = db.abs_faces if (db.stacey or db.pml is not None) else None
        pml = db.pml
        nsrc = self.sources.n if self.sources is not None else 0
        nrec = self.receivers.n if self.receivers is not None else 0
        if nsrc:
            sa = np.transpose(self.sources.sourcearrays, (1, 2, 0))  # (125, 3, ns): s fastest
            src_ispec = self.sources.ispec + 1
        else:
            sa = np.zeros(1, np.float32)
            src_ispec = np.zeros(1, np.int32)
        if nrec:
            R = self.receivers
            rec_ispec = R.ispec + 1
            hx, hy, hz = (R.hxir.astype(np.float32), R.hetar.astype(np.float32), R.hgammar.astype(np.float32))
            nu = np.transpose(R.nu, (0, 2, 1))
        else:
            rec_ispec = np.zeros(1, np.int32)
            hx = hy = hz = np.zeros(5, np.float32)
            nu = np.eye(3)[None]
        sv = lambda c: int(c in self.save)
        L.prepare_constants_device_(
            self.MP, A.i(5), A.i(db.nspec), A.i(db.nglob), A.i(nirr), A.ia(db.irregular_element_number),
            *[A.fa(j) for j in jac], A.f(db.xix_regular), A.f(db.jacobian_regular), A.ia(db.ibool.ravel() + 1),
            A.i(0), A.i(0), A.ia([0]), A.ia([0]),
            A.fa(g.hprime.ravel(order="F")), A.fa(g.hprimewgll.ravel(order="F")),
            A.fa(g.wgllwgll_xy.ravel(order="F")), A.fa(g.wgllwgll_xz.ravel(order="F")),
            A.fa(g.wgllwgll_yz.ravel(order="F")),
            A.i(int(db.stacey and db.abs_faces.n > 0)),
            A.ia(f.ispec + 1 if f is not None else [0]), A.ia(f.ijk.ravel() + 1 if f is not None else [0]),
            A.fa(f.normal.ravel() if f is not None else [0]), A.fa(f.jacobian2Dw.ravel() if f is not None else [0]),
            A.i(f.n if f is not None else 0),
            A.ia(np.ones(db.nspec, np.int32)),
            A.i(nsrc), A.i(nsrc), A.fa(sa.ravel()), A.ia(np.zeros(max(nsrc, 1), np.int32)), A.ia(src_ispec),
            A.ia(rec_ispec), A.i(nrec), A.i(nrec), A.i(self.sim), A.i(0),
            A.i(db.nspec_acoustic), A.i(db.nspec_elastic),
            A.ia(db.ispec_is_acoustic.astype(np.int32)), A.ia(db.ispec_is_elastic.astype(np.int32)),
            A.i(0), A.i(int(save_forward)), A.fa(hx.ravel()), A.fa(hy.ravel()), A.fa(hz.ravel()),
            A.da(nu.ravel()), A.da(np.zeros(9 * max(nsrc, 1))), A.ia(np.zeros(max(nrec, 1), np.int32)),
            A.i(self.nlength), A.i(sv("d")), A.i(sv("v")), A.i(sv("a")), A.i(sv("p")), A.i(1), A.i(0),
            A.i(int(self.undo_att)), A.i(int(pml is not None)))
        stage("prepare_constants_device")
        L.sf_set_options(self.MP, C.c_int(int(exact)), C.c_int(int(assembly)))
        for name, value in self.tuning.items():
            L.sf_set_tuning(self.MP, name.encode(), C.c_int(int(value)))
        if nirr:
            # the CPU solver's jacobianstore (prepare_constants_device itself recomputes it)
            L.sf_set_jacobian(self.MP, A.fa(db.jacobian.ravel()))
        nloc = db.nspec * 125
        if self.acoustic:
            n, ph = _fortran_phase_list(db.acoustic_elements())
            fr = db.free_faces
            cp = db.coupling_faces
            L.prepare_fields_acoustic_device_(
                self.MP, A.fa(db.rmass_acoustic), A.fa(db.rhostore.ravel()), A.fa(db.kappastore.ravel()), A.i(n),
                A.ia(ph), A.i(0), A.i(fr.n), A.ia(fr.ispec + 1), A.ia(fr.ijk.ravel() + 1), A.i(0), A.fa([0]),
                A.i(int(self.elastic)), A.i(cp.n), A.ia(cp.ispec + 1), A.ia(cp.ijk.ravel() + 1),
                A.fa(cp.normal.ravel()), A.fa(cp.jacobian2Dw.ravel()), A.i(0), A.i(0), A.ia([0]))
            if self.sim == 3:
                # reconstructed potentials and the acoustic kernel arrays
                L.prepare_fields_acoustic_adj_dev_(self.MP, A.i(0))
        stage("acoustic fields")
        if self.elastic:
            n, ph = _fortran_phase_list(db.elastic_elements())
            att = self.att
            has_att = att is not None
            strain = int(has_att or self.sim == 3)
            z = np.zeros(nloc if (strain or has_att) else 1, np.float32)
            if has_att:
                Rz = np.zeros(3 * nloc, np.float32)
                fc = att.factor_common.ravel()
                fck = att.factor_common_kappa.ravel()
                al, be, ga = att.alphaval, att.betaval, att.gammaval
                kap = att.kappastore.ravel()
                mu = att.mustore.ravel()
            else:
                Rz = np.zeros(1, np.float32)
                fc = fck = np.zeros(1, np.float32)
                al = be = ga = np.zeros(3, np.float32)
                kap = db.kappastore.ravel()
                mu = db.mustore.ravel()
            self.b_absorb_field = np.zeros(3 * 25 * max(db.abs_faces.n, 1), np.float32)
            L.prepare_fields_elastic_device_(
                self.MP, A.fa(db.rmassx), A.fa(db.rmassy), A.fa(db.rmassz), A.fa(db.rho_vp.ravel()),
                A.fa(db.rho_vs.ravel()), A.fa(kap), A.fa(mu), A.i(n), A.ia(ph),
                abi.ptr(self.b_absorb_field), A.i(4 * self.b_absorb_field.size), A.i(strain),
                A.fa(z), A.fa(z), A.fa(z), A.fa(z), A.fa(z), A.i(int(has_att)), A.i(3 * nloc if has_att else 1),
                A.fa(Rz), A.fa(Rz), A.fa(Rz), A.fa(Rz), A.fa(Rz), A.fa(fc), A.fa(Rz), A.fa(z), A.fa(fck),
                A.fa(al), A.fa(be), A.fa(ga), A.i(0), A.fa([0]), A.i(0), A.fa([0]), A.ia([0]), A.ia([0]), A.i(0),
                A.i(int(self.acoustic)), A.i(0), A.i(0), A.ia([0]), A.i(0), *[A.fa([0])] * 21)
            if self.sim == 3:
                # backward/reconstructed field, its strain and memory variables, kernel arrays
                zb = np.zeros(3 * nloc if has_att else 1, np.float32)
                L.prepare_fields_elastic_adj_dev_(
                    self.MP, A.i(3 * db.nglob), A.fa(z), A.fa(z), A.fa(z), A.fa(z), A.fa(z), A.fa(z), A.fa(z),
                    A.i(3 * nloc if has_att else 1), A.fa(zb), A.fa(zb), A.fa(zb), A.fa(zb), A.fa(zb), A.fa(zb),
                    A.fa(z), A.fa(al), A.fa(be), A.fa(ga), A.i(0), A.i(0))
                if nrec:
                    R = self.receivers
                    L.prepare_sim2_or_3_const_device_(
                        self.MP, A.i(nrec), A.i(self.nstep), A.fa(R.hxir.astype(np.float32).ravel()),
                        A.fa(R.hetar.astype(np.float32).ravel()), A.fa(R.hgammar.astype(np.float32).ravel()),
                        A.i(nrec), A.ia(np.zeros(nrec, np.int32)), A.ia(rec_ispec))
        if self.elastic and pml is not None:
            # prepare_gpu.f90:182-205
            g = get_gll()
            w = np.asarray(g.wgll, np.float64)
            wc = np.array([w[i] * w[j] * w[k] for k in range(5) for j in range(5) for i in range(5)], np.float64)
            L.prepare_fields_elastic_pml_(
                self.MP, A.i(pml.n), A.ia(pml.is_cpml(db.nspec).astype(np.int32)), A.ia(pml.cpml_to_spec + 1),
                A.ia(pml.spec_to_cpml(db.nspec)), A.fa(pml.coef_alpha.ravel()), A.fa(pml.coef_beta.ravel()),
                A.fa(pml.coef_abar.ravel()), A.fa(pml.coef_strain.ravel()), A.fa(wc.astype(np.float32)),
                A.fa(db.rhostore.ravel()))
        if self.elastic and self.element_order:
            if self.element_order == "morton":
                order = morton_order(db, db.elastic_elements())
            elif self.element_order.startswith("brick"):
                by, bz = (int(v) for v in self.element_order[5:].split("x"))
                order = brick_order(db, db.elastic_elements(), by, bz)
            else:
                raise ValueError(f"unknown element order {self.element_order!r}")
            L.sf_set_element_order(self.MP, C.c_int(len(order)), abi.ptr(order))
        stage("elastic fields")
        n3 = 3 * db.nglob
        self.displ = np.zeros(n3, np.float32)
        self.veloc = np.zeros(n3, np.float32)
        self.accel = np.zeros(n3, np.float32)
        if self.elastic:
            L.transfer_fields_el_to_device_(A.i(n3), abi.ptr(self.displ), abi.ptr(self.veloc), abi.ptr(self.accel),
                                            self.MP)
        if self.acoustic:
            self.pot = np.zeros(db.nglob, np.float32)
            L.transfer_fields_ac_to_device_(A.i(db.nglob), abi.ptr(self.pot), abi.ptr(self.pot), abi.ptr(self.pot),
                                            self.MP)
        nrec = max(nrec, 1)
        self.seis_d = np.zeros((self.nlength, nrec, 3), np.float32)
        self.seis_v = np.zeros((self.nlength, nrec, 3), np.float32)
        self.seis_a = np.zeros((self.nlength, nrec, 3), np.float32)
        self.seis_p = np.zeros((self.nlength, nrec, 3), np.float32)
        # constants (prepare_timerun.F90:389-441)
        dt = np.float32(db.dt)
        self.deltat = dt
        self.deltatover2 = np.float32(dt / np.float32(2))
        self.deltatsqover2 = np.float32(dt * dt / np.float32(2))
        self._c = {k: C.c_float(v) for k, v in dict(dt=self.deltat, dt2=self.deltatover2,
                                                  dtsq2=self.deltatsqover2).items()}
        self._i = {v: C.c_int(v) for v in (0, 1, 2, 3)}
        self.stf_table = self.sources.stf_table(self.nstep) if nsrc else None

    # ------------------------------------------------------------------------------ stepping
    def _ref(self, v):
        return C.byref(self._i[v])

    def step(self):
        """One forward time step in the reference's call order."""
        L, MP = self.L, self.MP
        self.it += 1
        it = self.it
        one, two, zero = self._ref(1), self._ref(2), self._ref(0)
        dt, dt2, dtsq2 = (C.byref(self._c[k]) for k in ("dt", "dt2", "dtsq2"))
        b_dt2 = C.byref(C.c_float(0.0))
        if self.acoustic:
            L.update_displacement_ac_cuda_(MP, dt, dt2, dtsq2, one)
        if self.elastic:
            L.update_displacement_cuda_(MP, dt, dt2, dtsq2, one)
        ncf = C.c_int(self.db.coupling_faces.n)
        if self.acoustic:
            L.acoustic_enforce_free_surf_cuda_(MP, zero, one)
            for iphase in (1, 2):
                ip = C.c_int(iphase)
                L.compute_forces_acoustic_cuda_(MP, C.byref(ip), zero, zero, one)
                if iphase == 1:
                    L.compute_stacey_acoustic_cuda_(MP, C.byref(ip), None, one)
                    if self.elastic:
                        L.compute_coupling_ac_el_cuda_(MP, C.byref(ip), C.byref(ncf), one)
                    if self.stf_table is not None:
                        # compute_add_sources_acoustic_GPU: sources in fluid elements (the library adds
                        # only those; the elastic ones come with the elastic forces)
                        stf = np.ascontiguousarray(self.stf_table[it - 1], np.float64)
                        L.compute_add_sources_ac_cuda_(MP, C.byref(C.c_int(len(stf))), stf.ctypes.data_as(C.c_void_p),
                                                       zero)
            L.kernel_3_acoustic_cuda_(MP, dt2, b_dt2, one)
            L.acoustic_enforce_free_surf_cuda_(MP, zero, one)
        if self.elastic:
            strain = C.c_int(int(self.att is not None))
            att = C.c_int(int(self.att is not None))
            for iphase in (1, 2):
                ip = C.c_int(iphase)
                L.compute_forces_viscoelastic_cuda_(MP, C.byref(ip), dt, zero, zero, C.byref(strain), C.byref(att), one)
                if iphase == 1:
                    L.compute_stacey_viscoelastic_cuda_(MP, C.byref(ip), abi.ptr(self.b_absorb_field), one)
                    if self.acoustic:
                        L.compute_coupling_el_ac_cuda_(MP, C.byref(ip), C.byref(ncf), one)
                    if self.stf_table is not None:
                        stf = np.ascontiguousarray(self.stf_table[it - 1], np.float64)
                        L.compute_add_sources_el_cuda_(MP, stf.ctypes.data_as(C.c_void_p), C.byref(C.c_int(len(stf))))
            if self.db.pml is not None:
                # Dirichlet condition on the outer edges of the C-PML layers
                L.pml_impose_boundary_condition_elastic_cuda_(MP)
            L.kernel_3_a_cuda_(MP, dt2, b_dt2, zero, one)
        if self.receivers is not None and self.receivers.n:
            self.seismo_current += 1
            L.compute_seismograms_cuda_(
                MP, abi.ptr(self.seis_d), abi.ptr(self.seis_v), abi.ptr(self.seis_a), abi.ptr(self.seis_p),
                C.byref(C.c_int(self.seismo_current)), C.byref(C.c_int(self.nlength)), C.byref(C.c_int(it)),
                C.byref(C.c_int(self.nstep)), C.byref(C.c_int(int(self.acoustic))),
                C.byref(C.c_int(int(self.elastic))), C.byref(C.c_int(int(self.pressure_trick))))

    def run(self, nstep: int | None = None):
        """Step ``nstep`` times; returns the first recorded kind of seismograms (nstep, nrec, 3)."""
        n = self.nstep if nstep is None else nstep
        for _ in range(n):
            self.step()
        return self.seismograms(self.save[0] if self.save else "d")

    def seismograms(self, kind: str = "d") -> np.ndarray:
        """Recorded seismograms of one kind ("d", "v", "a" or "p"; pressure in component 0)."""
        return {"d": self.seis_d, "v": self.seis_v, "a": self.seis_a, "p": self.seis_p}[kind][: self.seismo_current]

    def synchronize(self):
        self.L.sf_synchronize(self.MP)

    def fields(self):
        """(displ, veloc, accel) copied back from the device."""
        A = abi.Args()
        n3 = 3 * self.db.nglob
        self.L.transfer_fields_el_from_device_(A.i(n3), abi.ptr(self.displ), abi.ptr(self.veloc),
                                               abi.ptr(self.accel), self.MP)
        return self.displ.copy(), self.veloc.copy(), self.accel.copy()

    def potentials(self):
        """(potential, first and second time derivatives) of the fluid copied back from the device."""
        A = abi.Args()
        n = self.db.nglob
        out = [np.zeros(n, np.float32) for _ in range(3)]
        self.L.transfer_fields_ac_from_device_(A.i(n), abi.ptr(out[0]), abi.ptr(out[1]), abi.ptr(out[2]), self.MP)
        return tuple(out)

    def max_norm(self) -> float:
        out = C.c_float(0)
        self.L.get_norm_elastic_from_device_(C.byref(out), self.MP, self._ref(1))
        return float(out.value)

    def close(self):
        if self.mp.value:
            z = C.c_int(0)
            self.L.prepare_cleanup_device_(self.MP, C.byref(z), C.byref(z), C.byref(z))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
