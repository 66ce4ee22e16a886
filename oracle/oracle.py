"""CPU oracle for the single-partition SPECFEM3D solver loop (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import this module, and only as the checker / CPU baseline. The product path never
calls it.

``OracleSolver`` drives the C restatement in ``specfem_oracle.c`` through one time
step in exactly the order of the reference host loop at NPROC=1:

* ``iterate_time.F90:219-330``: Newmark predictor (acoustic, then elastic), acoustic
  forces, elastic forces, seismograms, kernels;
* ``compute_forces_acoustic_calling_routine.f90:54-343``: free surface, Stacey,
  coupling (uses the predicted displacement), sources, element loop (all elements are
  "inner" at NPROC=1, so the element loop runs after the boundary terms), mass +
  corrector, free surface again;
* ``compute_forces_viscoelastic_calling_routine.F90:30-435``: Stacey, coupling
  (uses the updated fluid potential), sources, element loop, mass + corrector.

Precision: ``"f32"`` is the FP32 oracle (same operation order as the reference's
CUSTOM_REAL=4 CPU build, no fused multiply-add); ``"f64"`` is the FP64 oracle on the
same (float32) database values.

Parity status of each restated function is recorded in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build_oracle(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "specfem_oracle.c")
    if force or not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return path


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(build_oracle())
    return _LIB


_LIB_OMP = None


def lib_omp():
    """The multi-threaded FP32 build (element loops in parallel, atomic scatter: the summation
    order at shared points varies), for the CPU baseline only; parity checks use ``lib()``."""
    global _LIB_OMP
    if _LIB_OMP is None:
        build_oracle()
        path = os.path.join(_HERE, "liboracle_omp.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        _LIB_OMP = C.CDLL(path)
    return _LIB_OMP


def get_gll_weights():
    from specfem3d_amd.gll import get_gll
    return np.asarray(get_gll().wgll, np.float64)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class AttenuationState:
    alphaval: np.ndarray
    betaval: np.ndarray
    gammaval: np.ndarray
    factor_common: np.ndarray        # (nspec*125*3)
    factor_common_kappa: np.ndarray


class OracleSolver:
    """Forward (and adjoint-with-stored-forward) time stepping on the CPU."""

    def __init__(self, db, sources=None, receivers=None, precision: str = "f32", attenuation=None,
                 compute_strain: bool = False, store_eps_trace3: bool = False, threads: int = 1):
        """``threads`` > 1 runs the multi-threaded FP32 build (CPU baseline timing only)."""
        self.db = db
        self.prec = precision
        self.dt_np = np.float32 if precision == "f32" else np.float64
        self.sfx = "_f32" if precision == "f32" else "_f64"
        if threads > 1:
            if precision != "f32":
                raise ValueError("the multi-threaded oracle is built for FP32 only")
            self.L = lib_omp()
            self.L.orc_set_threads_f32(C.c_int(int(threads)))
        else:
            self.L = lib()
        R = self.dt_np
        from specfem3d_amd.gll import get_gll
        g = get_gll()
        self.hprime = np.ascontiguousarray(g.hprime, dtype=R)
        self.hprimewgll = np.ascontiguousarray(g.hprimewgll, dtype=R)
        self.wgllwgll = np.ascontiguousarray(g.wgllwgll_xy, dtype=R)
        self.ibool = np.ascontiguousarray(db.ibool.reshape(-1), dtype=np.int32)
        self.irr = np.ascontiguousarray(db.irregular_element_number, dtype=np.int32)
        self.jac9 = np.ascontiguousarray(db.jac9.reshape(-1), dtype=R)
        self.jac = np.ascontiguousarray(db.jacobian.reshape(-1), dtype=R)
        self.kappa = np.ascontiguousarray(db.kappastore.reshape(-1), dtype=R)
        self.kappa_ac = self.kappa    # kappastore of the fluid source terms (not rescaled by attenuation)
        self.mu = np.ascontiguousarray(db.mustore.reshape(-1), dtype=R)
        self.rho = np.ascontiguousarray(db.rhostore.reshape(-1), dtype=R)
        self.rho_vp = np.ascontiguousarray(db.rho_vp.reshape(-1), dtype=R)
        self.rho_vs = np.ascontiguousarray(db.rho_vs.reshape(-1), dtype=R)
        self.rmassx = db.rmassx.astype(R)
        self.rmassy = db.rmassy.astype(R)
        self.rmassz = db.rmassz.astype(R)
        self.rmass_ac = db.rmass_acoustic.astype(R)
        self.is_el = np.ascontiguousarray(db.ispec_is_elastic.astype(np.uint8))
        self.is_ac = np.ascontiguousarray(db.ispec_is_acoustic.astype(np.uint8))
        self.el_elems = db.elastic_elements()
        self.ac_elems = db.acoustic_elements()
        self.elastic = len(self.el_elems) > 0
        self.acoustic = len(self.ac_elems) > 0
        f = db.abs_faces
        self.abs = (f.n, np.ascontiguousarray(f.ispec, np.int32), np.ascontiguousarray(f.ijk, np.int32),
                    np.ascontiguousarray(f.normal, R), np.ascontiguousarray(f.jacobian2Dw, R))
        f = db.free_faces
        self.free = (f.n, np.ascontiguousarray(f.ispec, np.int32), np.ascontiguousarray(f.ijk, np.int32))
        f = db.coupling_faces
        self.cpl = (f.n, np.ascontiguousarray(f.ispec, np.int32), np.ascontiguousarray(f.ijk, np.int32),
                    np.ascontiguousarray(f.normal, R), np.ascontiguousarray(f.jacobian2Dw, R))
        self.stacey = db.stacey and db.abs_faces.n > 0
        dt = db.dt
        self.deltat = R(np.float32(dt)) if precision == "f32" else R(dt)
        # reference: deltat = real(DT); deltatover2 = deltat/2; deltatsqover2 = deltat*deltat/2
        self.deltatover2 = R(self.deltat / R(2))
        self.deltatsqover2 = R(self.deltat * self.deltat / R(2))
        n = db.nglob
        self.displ = np.zeros(3 * n, R)
        self.veloc = np.zeros(3 * n, R)
        self.accel = np.zeros(3 * n, R)
        self.pot = np.zeros(n, R)
        self.potd = np.zeros(n, R)
        self.potdd = np.zeros(n, R)
        self.sources = sources
        self.receivers = receivers
        if sources is not None and sources.n:
            self.src_ispec = np.ascontiguousarray(sources.ispec, np.int32)
            self.src_arr = np.ascontiguousarray(sources.sourcearrays.reshape(-1), R)
        if receivers is not None and receivers.n:
            self.rec_ispec = np.ascontiguousarray(receivers.ispec, np.int32)
            self.hx = np.ascontiguousarray(receivers.hxir, np.float64)
            self.hy = np.ascontiguousarray(receivers.hetar, np.float64)
            self.hz = np.ascontiguousarray(receivers.hgammar, np.float64)
            self.nu = np.ascontiguousarray(receivers.nu, np.float64)
        self.att = attenuation
        self.compute_strain = compute_strain or (attenuation is not None)
        ne = db.nspec * 125
        if self.compute_strain:
            self.eps = [np.zeros(ne, R) for _ in range(6)]  # xx,yy,xy,xz,yz,trace
        else:
            self.eps = [None] * 6
        self.eps_tr3 = np.zeros(ne, R) if store_eps_trace3 else None
        if attenuation is not None:
            # moduli scaled to the attenuation reference frequency (prepare_attenuation.f90)
            self.kappa = np.ascontiguousarray(attenuation.kappastore.reshape(-1), dtype=R)
            self.mu = np.ascontiguousarray(attenuation.mustore.reshape(-1), dtype=R)
            self.R = [np.zeros(ne * 3, R) for _ in range(6)]  # xx,yy,xy,xz,yz,trace
            self.alpha = attenuation.alphaval.astype(R)
            self.beta = attenuation.betaval.astype(R)
            self.gamma = attenuation.gammaval.astype(R)
            self.fc = np.ascontiguousarray(attenuation.factor_common.reshape(-1), R)
            self.fck = np.ascontiguousarray(attenuation.factor_common_kappa.reshape(-1), R)
        self.pml = db.pml
        if self.pml is not None:
            self._setup_pml()
        self.it = 0
        self.adjoint_sign = 0   # 1 for the adjoint wavefield of a kernel run (coupling sign flip)
        # tests plant a sign mistake in one coupling direction through these (never set otherwise)
        self.plant_flip_ac_el = 0
        self.plant_flip_el_ac = 0
        self.stf_table = None
        self.absorb_record = None   # SAVE_FORWARD: Stacey terms of the step (3*25 per face), or None

    def _setup_pml(self):
        """C-PML state (pml_allocate_arrays.f90: zero memory variables) and the element split of
        compute_forces_viscoelastic (non-PML elements) / compute_forces_viscoelastic_PML (PML ones)."""
        db, pml, R = self.db, self.pml, self.dt_np
        n = pml.n
        self.spec_to_cpml = np.ascontiguousarray(pml.spec_to_cpml(db.nspec), np.int32)
        self.cpml_to_spec = np.ascontiguousarray(pml.cpml_to_spec, np.int32)
        is_pml = pml.is_cpml(db.nspec)
        self.el_elems_all = self.el_elems
        self.el_elems = np.ascontiguousarray(self.el_elems_all[~is_pml[self.el_elems_all]], np.int32)
        self.pml_old = np.zeros(3 * 125 * n, R)
        self.pml_new = np.zeros(3 * 125 * n, R)
        self.rm3 = [np.zeros(3 * 125 * n, R) for _ in range(9)]
        self.rm1 = [np.zeros(125 * n, R) for _ in range(12)]
        self.rdisp = np.zeros(9 * 125 * n, R)
        self._rm3_ptrs = (C.c_void_p * 9)(*[a.ctypes.data for a in self.rm3])
        self._rm1_ptrs = (C.c_void_p * 12)(*[a.ctypes.data for a in self.rm1])
        self.coef_alpha = np.ascontiguousarray(pml.coef_alpha, R).reshape(-1)
        self.coef_beta = np.ascontiguousarray(pml.coef_beta, R).reshape(-1)
        self.coef_abar = np.ascontiguousarray(pml.coef_abar, R).reshape(-1)
        self.coef_strain = np.ascontiguousarray(pml.coef_strain, R).reshape(-1)
        w = get_gll_weights()
        # wgll_cube = sngl(wxgll(i) wygll(j) wzgll(k)) (prepare_gravity.f90:61-67)
        wc = np.empty(125)
        for k in range(5):
            for j in range(5):
                for i in range(5):
                    wc[i + 5 * j + 25 * k] = w[i] * w[j] * w[k]
        self.wgll_cube = np.ascontiguousarray(wc.astype(np.float32), R)

    def _pml_store(self, dest):
        self._call("orc_pml_store", C.c_int(self.pml.n), _p(self.cpml_to_spec), _p(self.ibool), _p(self.displ),
                   _p(self.veloc), _p(self.accel), self._r(self.deltatover2), self._r(self.deltatsqover2), _p(dest))

    def _pml_forces(self):
        db = self.db
        self._call("orc_forces_elastic_pml", C.c_int(len(self.el_elems_all)), _p(self.el_elems_all),
                   _p(self.spec_to_cpml), C.c_int(self.pml.n), _p(self.ibool), _p(self.irr),
                   self._r(db.xix_regular), self._r(db.jacobian_regular), _p(self.jac9), _p(self.jac),
                   _p(self.hprime), _p(self.hprimewgll), _p(self.wgllwgll), _p(self.wgll_cube), _p(self.kappa),
                   _p(self.mu), _p(self.rho), _p(self.displ), _p(self.veloc), _p(self.accel), _p(self.pml_old),
                   _p(self.pml_new), self._rm3_ptrs, self._rm1_ptrs, _p(self.rdisp), _p(self.coef_alpha),
                   _p(self.coef_beta), _p(self.coef_abar), _p(self.coef_strain))

    def _pml_dirichlet(self):
        n, isp, ijk, nrm, jw = self.abs
        self._call("orc_pml_dirichlet", C.c_int(n), _p(isp), _p(ijk), _p(self.ibool), _p(self.is_el),
                   _p(self.spec_to_cpml), _p(self.displ), _p(self.veloc), _p(self.accel), _p(self.pml_old),
                   _p(self.pml_new))

    # -- pieces ---------------------------------------------------------------------
    def _call(self, name, *args):
        return getattr(self.L, name + self.sfx)(*args)

    def predictor(self):
        R = self.dt_np
        if self.acoustic:
            self._call("orc_update_displ", C.c_int64(self.db.nglob), _p(self.pot), _p(self.potd), _p(self.potdd),
                       self._r(self.deltat), self._r(self.deltatover2), self._r(self.deltatsqover2))
        if self.elastic:
            if self.pml is not None:
                self._pml_store(self.pml_old)   # update_displacement_elastic: PML stores the old field
            self._call("orc_update_displ", C.c_int64(3 * self.db.nglob), _p(self.displ), _p(self.veloc),
                       _p(self.accel), self._r(self.deltat), self._r(self.deltatover2), self._r(self.deltatsqover2))
            if self.pml is not None:
                self._pml_store(self.pml_new)   # ... and the updated one

    def _r(self, v):
        return C.c_float(v) if self.prec == "f32" else C.c_double(v)

    def forces_elastic(self):
        db = self.db
        att = self.att is not None
        e = self.eps
        if att:
            Rr = self.R
            attargs = (_p(self.alpha), _p(self.beta), _p(self.gamma), _p(self.fc), _p(self.fck),
                       _p(Rr[0]), _p(Rr[1]), _p(Rr[2]), _p(Rr[3]), _p(Rr[4]), _p(Rr[5]))
        else:
            attargs = (None,) * 11
        self._call("orc_forces_elastic", C.c_int(len(self.el_elems)), _p(self.el_elems), _p(self.ibool),
                   _p(self.irr), self._r(db.xix_regular), self._r(db.jacobian_regular), _p(self.jac9),
                   _p(self.jac), _p(self.hprime), _p(self.hprimewgll), _p(self.wgllwgll), _p(self.kappa),
                   _p(self.mu), _p(self.displ), _p(self.veloc), _p(self.accel), self._r(self.deltat),
                   C.c_int(int(att)), C.c_int(int(self.compute_strain)), *attargs,
                   _p(e[0]), _p(e[1]), _p(e[2]), _p(e[3]), _p(e[4]), _p(e[5]), _p(self.eps_tr3))
        if self.pml is not None:
            self._pml_forces()

    def forces_acoustic(self):
        db = self.db
        self._call("orc_forces_acoustic", C.c_int(len(self.ac_elems)), _p(self.ac_elems), _p(self.ibool),
                   _p(self.irr), self._r(db.xix_regular), self._r(db.jacobian_regular), _p(self.jac9),
                   _p(self.jac), _p(self.hprime), _p(self.hprimewgll), _p(self.wgllwgll), _p(self.rho),
                   _p(self.pot), _p(self.potdd))

    def stf_values(self, it):
        s = self.sources
        if self.stf_table is not None:
            v = self.stf_table[it - 1]
        else:
            v = s.stf(it)
        return np.asarray(v, np.float64).astype(np.float32).astype(self.dt_np)  # stf_used = real(stf)

    def step(self, adjoint_source=None):
        """One forward time step (1-based step index self.it+1)."""
        self.it += 1
        it = self.it
        self.predictor()
        if self.acoustic:
            nf, fis, fijk = self.free
            if nf:
                self._call("orc_free_surface", C.c_int(nf), _p(fis), _p(fijk), _p(self.ibool), _p(self.is_ac),
                           _p(self.pot), _p(self.potd), _p(self.potdd))
            if self.stacey:
                n, isp, ijk, nrm, jw = self.abs
                self._call("orc_stacey_acoustic", C.c_int(n), _p(isp), _p(ijk), _p(jw), _p(self.ibool),
                           _p(self.is_ac), _p(self.rho), _p(self.kappa), _p(self.potd), _p(self.potdd))
            if self.elastic and self.cpl[0]:
                n, isp, ijk, nrm, jw = self.cpl
                self._call("orc_coupling_ac_el", C.c_int(n), _p(isp), _p(ijk), _p(nrm), _p(jw), _p(self.ibool),
                           _p(self.displ), _p(self.potdd), C.c_int(self.adjoint_sign ^ self.plant_flip_ac_el))
            # sources and adjoint sources in fluid elements (compute_forces_acoustic_calling_routine.f90:
            # after Stacey and coupling, before the stiffness of phase 2)
            if self.sources is not None and self.sources.n:
                stf = np.ascontiguousarray(self.stf_values(it), self.dt_np)
                self._call("orc_add_sources_ac", C.c_int(self.sources.n), _p(self.src_ispec), _p(self.src_arr),
                           _p(stf), _p(self.ibool), _p(self.is_ac), _p(self.kappa_ac), _p(self.potdd))
            if adjoint_source is not None:
                adj = np.ascontiguousarray(adjoint_source, self.dt_np)
                self._call("orc_add_adjoint_sources_ac", C.c_int(self.receivers.n), _p(self.rec_ispec),
                           _p(self.hx), _p(self.hy), _p(self.hz), _p(adj), _p(self.ibool), _p(self.is_ac),
                           _p(self.potdd))
            self.forces_acoustic()
            self._call("orc_kernel3_acoustic", C.c_int64(self.db.nglob), _p(self.potd), _p(self.potdd),
                       _p(self.rmass_ac), self._r(self.deltatover2))
            nf, fis, fijk = self.free
            if nf:
                self._call("orc_free_surface", C.c_int(nf), _p(fis), _p(fijk), _p(self.ibool), _p(self.is_ac),
                           _p(self.pot), _p(self.potd), _p(self.potdd))
        if self.elastic:
            if self.stacey:
                n, isp, ijk, nrm, jw = self.abs
                self._call("orc_stacey_elastic", C.c_int(n), _p(isp), _p(ijk), _p(nrm), _p(jw), _p(self.ibool),
                           _p(self.is_el), _p(self.rho_vp), _p(self.rho_vs), _p(self.veloc), _p(self.accel),
                           None if self.absorb_record is None else _p(self.absorb_record))
            if self.acoustic and self.cpl[0]:
                n, isp, ijk, nrm, jw = self.cpl
                self._call("orc_coupling_el_ac", C.c_int(n), _p(isp), _p(ijk), _p(nrm), _p(jw), _p(self.ibool),
                           _p(self.potdd), _p(self.accel), C.c_int(self.adjoint_sign ^ self.plant_flip_el_ac))
            if self.sources is not None and self.sources.n:
                stf = np.ascontiguousarray(self.stf_values(it), self.dt_np)
                self._call("orc_add_sources_el", C.c_int(self.sources.n), _p(self.src_ispec), _p(self.src_arr),
                           _p(stf), _p(self.ibool), _p(self.is_el), _p(self.accel))
            if adjoint_source is not None:
                adj = np.ascontiguousarray(adjoint_source, self.dt_np)
                self._call("orc_add_adjoint_sources_el", C.c_int(self.receivers.n), _p(self.rec_ispec),
                           _p(self.hx), _p(self.hy), _p(self.hz), _p(adj), _p(self.ibool), _p(self.accel),
                           _p(self.is_el))
            self.forces_elastic()
            if self.pml is not None:
                self._pml_dirichlet()   # compute_forces_viscoelastic_calling_routine.F90:361-365
            self._call("orc_kernel3_elastic", C.c_int64(self.db.nglob), _p(self.veloc), _p(self.accel),
                       _p(self.rmassx), _p(self.rmassy), _p(self.rmassz), self._r(self.deltatover2))

    def backward_step(self, record: np.ndarray, stf_index: int):
        """One step of the time-reversed reconstruction (elastic domains), iterate_time.F90:233-262 with
        b_deltat = -DT (prepare_timerun.F90:424-429; set by ``reversed_solver``): predictor, then the
        Stacey record ``record`` subtracted in place of Stacey (compute_stacey_viscoelastic.F90:212-287),
        the sources at stf_table[stf_index], stiffness, mass and corrector. The reference passes the
        record of forward step NSTEP-it+1 and the source time (NSTEP-it) DT - t0, i.e. stf_index =
        NSTEP-it (compute_add_sources_viscoelastic.F90:650-690)."""
        if self.pml is not None:
            raise ValueError("C-PML runs are forward only here (the reference reconstructs PML fields from "
                             "stored interface values, which this restatement does not cover)")
        self.it += 1
        self.predictor()
        n, isp, ijk, nrm, jw = self.abs
        if self.stacey:
            rec = np.ascontiguousarray(record, self.dt_np)
            self._call("orc_stacey_elastic_replay", C.c_int(n), _p(isp), _p(ijk), _p(self.ibool), _p(self.is_el),
                       _p(rec), _p(self.accel))
        if self.sources is not None and self.sources.n:
            stf = np.ascontiguousarray(
                np.asarray(self.stf_table[stf_index], np.float64).astype(np.float32).astype(self.dt_np))
            self._call("orc_add_sources_el", C.c_int(self.sources.n), _p(self.src_ispec), _p(self.src_arr),
                       _p(stf), _p(self.ibool), _p(self.is_el), _p(self.accel))
        self.forces_elastic()
        self._call("orc_kernel3_elastic", C.c_int64(self.db.nglob), _p(self.veloc), _p(self.accel),
                   _p(self.rmassx), _p(self.rmassy), _p(self.rmassz), self._r(self.deltatover2))

    def seismogram(self, field: str = "displ") -> np.ndarray:
        """One sample of every receiver: ``field`` in "displ", "veloc", "accel" (vectors; in fluid
        elements (1/rho) grad of the potential of the same order), "pressure" (-potential_dot_dot in fluid
        elements) or "pressure_trick" (-potential, the source STF being the second derivative)."""
        out = np.zeros(3 * self.receivers.n, self.dt_np)
        if field in ("displ", "veloc", "accel") and self.elastic:
            f = {"displ": self.displ, "veloc": self.veloc, "accel": self.accel}[field]
            self._call("orc_seismo_elastic", C.c_int(self.receivers.n), _p(self.rec_ispec), _p(self.hx),
                       _p(self.hy), _p(self.hz), _p(self.nu), _p(self.ibool), _p(f), _p(out), _p(self.is_el))
        if self.acoustic:
            pot = {"displ": self.pot, "veloc": self.potd, "accel": self.potdd, "pressure": self.potdd,
                   "pressure_trick": self.pot}[field]
            self._call("orc_seismo_acoustic", C.c_int(self.receivers.n), _p(self.rec_ispec), _p(self.hx),
                       _p(self.hy), _p(self.hz), _p(self.nu), _p(self.ibool), _p(self.is_ac), _p(self.irr),
                       self._r(self.db.xix_regular), _p(self.jac9), _p(self.hprime), _p(self.rho), _p(pot),
                       C.c_int(int(field.startswith("pressure"))), _p(out))
        return out.reshape(-1, 3)

    def run(self, nstep: int, record: str = "displ", callback=None) -> np.ndarray:
        """Run ``nstep`` steps; returns seismograms (nstep, nrec, 3)."""
        nrec = self.receivers.n if self.receivers is not None else 0
        seis = np.zeros((nstep, nrec, 3), self.dt_np)
        if self.sources is not None and self.sources.n:
            self.stf_table = self.sources.stf_table(self.it + nstep) if self.stf_table is None else self.stf_table
        for s in range(nstep):
            self.step()
            if nrec:
                seis[s] = self.seismogram(record)
            if callback is not None:
                callback(self)
        return seis

    def max_norm(self) -> float:
        return float(self._call_ret("orc_max_norm", self.displ))

    def _call_ret(self, name, f):
        fn = getattr(self.L, name + self.sfx)
        fn.restype = C.c_float if self.prec == "f32" else C.c_double
        return fn(C.c_int64(self.db.nglob), _p(f))


def reversed_solver(db, sources, receivers, nstep: int, last, precision: str = "f32") -> OracleSolver:
    """The backward field of the classic kernel run: starts from the forward run's last frame
    ``last`` = (displ, veloc, accel) (read_forward_arrays.f90:29-126) and steps with b_deltat = -DT,
    b_deltatover2 = b_deltat/2, b_deltatsqover2 = b_deltat^2/2 (prepare_timerun.F90:424-429)."""
    b = OracleSolver(db, sources, receivers, precision)
    R = b.dt_np
    b.deltat = R(-b.deltat)
    b.deltatover2 = R(b.deltat / R(2))
    b.deltatsqover2 = R(b.deltat * b.deltat / R(2))
    b.stf_table = sources.stf_table(nstep) if sources is not None and sources.n else None
    for dst, src in zip((b.displ, b.veloc, b.accel), last):
        dst[:] = src
    return b


def element_strain(db, displ, precision="f32"):
    """Deviatoric strain and trace/3 of a displacement field (``compute_element_strain``)."""
    R = np.float32 if precision == "f32" else np.float64
    sfx = "_f32" if precision == "f32" else "_f64"
    from specfem3d_amd.gll import get_gll
    g = get_gll()
    L = lib()
    ne = db.nspec * 125
    out = [np.zeros(ne, R) for _ in range(6)]
    elems = db.elastic_elements()
    fn = getattr(L, "orc_element_strain" + sfx)
    rarg = (C.c_float if precision == "f32" else C.c_double)(db.xix_regular)
    fn(C.c_int(len(elems)), _p(elems), _p(np.ascontiguousarray(db.ibool.reshape(-1), np.int32)),
       _p(np.ascontiguousarray(db.irregular_element_number, np.int32)), rarg,
       _p(np.ascontiguousarray(db.jac9.reshape(-1), R)), _p(np.ascontiguousarray(g.hprime, R)),
       _p(np.ascontiguousarray(displ, R)), *[_p(o) for o in out])
    return out  # xx, yy, xy, xz, yz, trace/3


def _snapshot(s: OracleSolver) -> dict:
    """State saved per time subset (save_forward_arrays_undoatt, save_forward_arrays.f90:102-179)."""
    snap = {"it": s.it, "displ": s.displ.copy(), "veloc": s.veloc.copy(), "accel": s.accel.copy(),
            "pot": s.pot.copy(), "potd": s.potd.copy(), "potdd": s.potdd.copy()}
    if s.att is not None:
        snap["R"] = [r.copy() for r in s.R]
    return snap


def _restore(s: OracleSolver, snap: dict):
    """read_forward_arrays_undoatt (read_forward_arrays.f90:133-229) followed by the strain
    recomputation from the displacement (compute_strain_att_backward,
    compute_element_strain.F90:228-254)."""
    s.it = snap["it"]
    s.displ[:] = snap["displ"]
    s.veloc[:] = snap["veloc"]
    s.accel[:] = snap["accel"]
    s.pot[:] = snap["pot"]
    s.potd[:] = snap["potd"]
    s.potdd[:] = snap["potdd"]
    if s.att is not None:
        for r, v in zip(s.R, snap["R"]):
            r[:] = v
    _recompute_strain(s)


def _recompute_strain(s: OracleSolver):
    if not s.compute_strain:
        return
    e = element_strain(s.db, s.displ, s.prec)
    for c in range(5):
        s.eps[c][:] = e[c]
    if s.att is not None:
        s.eps[5][:] = s.dt_np(3) * e[5]
    if s.eps_tr3 is not None:
        s.eps_tr3[:] = e[5]


def kernel_run(db, sources, receivers, nstep: int, adj: np.ndarray, nt_dump: int, precision: str = "f64",
               attenuation=None, with_acoustic: bool = False, acoustic_adjoint: str = "potential"):
    """Checkpointed sensitivity-kernel run at NPROC=1 (elastic domains).

    Follows iterate_time_undoatt.F90:29-632 for SIMULATION_TYPE=1 with SAVE_FORWARD (snapshot at the
    start of every subset of ``nt_dump`` steps) and SIMULATION_TYPE=3 (subsets in reverse: restore a
    snapshot, recompute the subset storing b_displ per step, then step the adjoint field through the
    subset with b_displ read back time-reversed, :420-560). The adjoint source of step ``it`` is
    ``adj[NSTEP - it]`` (forward-time index NSTEP-it+1 of the .adj record, added while it < NSTEP;
    compute_add_sources_viscoelastic.F90:181-262 and compute_arrays_source.f90:203-291). Kernels
    accumulate after every adjoint step (compute_kernels.f90:67-222) with the reconstructed strain
    recomputed from b_displ.

    In fluid elements the buffered fields are b_potential and b_potential_dot_dot and the
    acoustic kernels accumulate as compute_kernels_ac (compute_kernels.f90:257-337), pairing the
    adjoint ``potential_acoustic`` with them (``acoustic_adjoint="potential"``, :269 and :325).
    ``acoustic_adjoint="potential_dot_dot"`` pairs the adjoint potential_dot_dot instead: the
    adjoint potential is driven by the adjoint displacement through the coupling
    (compute_coupling_acoustic_el.f90:105-110), so the misfit gradient needs its second time
    derivative.

    Returns (forward seismograms (nstep, nrec, 3), rho_kl, mu_kl, kappa_kl) -- the raw time
    integrals before save_adjoint_kernels' scaling -- plus (rho_ac_kl, kappa_ac_kl) when
    ``with_acoustic``.
    """
    if acoustic_adjoint not in ("potential", "potential_dot_dot"):
        raise ValueError("acoustic_adjoint must be 'potential' or 'potential_dot_dot'")
    nt_dump = min(nt_dump, nstep)
    fwd = OracleSolver(db, sources, receivers, precision, attenuation=attenuation)
    fwd.stf_table = sources.stf_table(nstep)
    nsub = -(-nstep // nt_dump)
    snaps = []
    seis = np.zeros((nstep, receivers.n, 3), fwd.dt_np)
    for s in range(nsub):
        snaps.append(_snapshot(fwd))
        n_s = nt_dump if s < nsub - 1 else nstep - (nsub - 1) * nt_dump
        for _ in range(n_s):
            fwd.step()
            seis[fwd.it - 1] = fwd.seismogram()
    del fwd

    R = np.float32 if precision == "f32" else np.float64
    b = OracleSolver(db, sources, receivers, precision, attenuation=attenuation)
    b.stf_table = sources.stf_table(nstep)
    adjs = OracleSolver(db, None, receivers, precision, attenuation=attenuation, compute_strain=True,
                        store_eps_trace3=True)
    adjs.adjoint_sign = 1
    ne = db.nspec * 125
    rho_kl, mu_kl, kappa_kl = np.zeros(ne, R), np.zeros(ne, R), np.zeros(ne, R)
    L = lib()
    fn = getattr(L, "orc_compute_kernels_el" + b.sfx)
    elems = db.elastic_elements()
    ib = np.ascontiguousarray(db.ibool.reshape(-1), np.int32)
    adj = np.ascontiguousarray(adj, R)
    ac_elems = db.acoustic_elements()
    rho_ac_kl, kappa_ac_kl = np.zeros(ne, R), np.zeros(ne, R)
    fna = getattr(L, "orc_compute_kernels_ac" + b.sfx)
    g = None
    if len(ac_elems):
        from specfem3d_amd.gll import get_gll
        g = get_gll()
        hp = np.ascontiguousarray(g.hprime, R)
        irr = np.ascontiguousarray(db.irregular_element_number, np.int32)
        jac9 = np.ascontiguousarray(db.jac9.reshape(-1), R)
        rho = np.ascontiguousarray(db.rhostore.reshape(-1), R)
    it = 0
    for i in range(1, nsub + 1):
        slot = nsub - i
        _restore(b, snaps[slot])
        n_i = nstep - (nsub - 1) * nt_dump if i == 1 else nt_dump
        buf = []
        for _ in range(n_i):
            b.step()
            buf.append((b.displ.copy(), b.pot.copy(), b.potdd.copy()))
        _recompute_strain(adjs)          # compute_strain_att before the adjoint subset
        for m in range(n_i):
            it += 1
            bd, bp, bpdd = buf[n_i - 1 - m]
            adjs.step(adjoint_source=adj[nstep - it] if it < nstep else None)
            if len(elems):
                be = element_strain(db, bd, precision)
                e = adjs.eps
                fn(C.c_int(len(elems)), _p(elems), _p(ib), _p(adjs.accel), _p(bd), _p(e[0]), _p(e[1]), _p(e[2]),
                   _p(e[3]), _p(e[4]), _p(adjs.eps_tr3), _p(be[0]), _p(be[1]), _p(be[2]), _p(be[3]), _p(be[4]),
                   _p(be[5]), adjs._r(adjs.deltat), _p(rho_kl), _p(mu_kl), _p(kappa_kl))
            if len(ac_elems):
                chi = adjs.potdd if acoustic_adjoint == "potential_dot_dot" else adjs.pot
                fna(C.c_int(len(ac_elems)), _p(ac_elems), _p(ib), _p(irr), adjs._r(db.xix_regular), _p(jac9), _p(hp),
                    _p(rho), _p(chi), _p(bp), _p(bpdd), adjs._r(adjs.deltat), _p(rho_ac_kl), _p(kappa_ac_kl))
    if with_acoustic:
        return seis, rho_kl, mu_kl, kappa_kl, rho_ac_kl, kappa_ac_kl
    return seis, rho_kl, mu_kl, kappa_kl
