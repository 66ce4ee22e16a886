"""Batched many-shot forward runs on one mesh (the ``sf_batch_*`` entry points of the C-ABI).

The reference runs one shot per solver process; a survey with many sources re-reads the mesh and
re-runs the whole loop per shot. ``BatchSolver`` prepares the mesh once through the reference ABI
(``GPUSolver`` with no sources or receivers), then steps all shots together in shot-interleaved
wavefields (``csrc/batch.hip``). The shots have their own sources, receivers and record lengths.
Shots-per-workgroup and the assembly mode come from the library's autotuner, and time steps are
replayed from captured hipGraphs. Each shot's seismograms are bit-identical to a single-shot
``GPUSolver`` run with the same options. Meshes with fluid elements (coupled to the solid) step each
shot's potentials too; sources and receivers lie in the solid.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .solver import GPUSolver


class BatchSolver:
    def __init__(self, db, shots, exact: bool = False, device: int = 0, shots_per_group: int | None = None,
                 assembly: int | None = None, attenuation=None):
        """``shots``: list of (SourceSet, ReceiverSet, nstep), one per shot. ``assembly``: 0 atomic
        scatter, 2 one launch per element colour, 3 coloured with the Newmark update fused into the
        colour launches (same bits as 2); None keeps the library default until ``autotune``.
        ``attenuation``: 3-SLS set-up of the mesh (each shot then carries its own memory variables;
        coloured assembly only)."""
        if not shots:
            raise ValueError("no shots")
        self.db = db
        self.shots = list(shots)
        self.nstep = [int(n) for _, _, n in self.shots]
        # Morton launch order: elements of a colour that sit close together share L2 lines of rows
        self.base = GPUSolver(db, None, None, nstep=max(self.nstep), exact=exact, assembly=0, device=device,
                              element_order="morton", attenuation=attenuation)
        L = self.L = self.base.L
        L.sf_batch_create.restype = C.c_long
        L.sf_batch_autotune.restype = C.c_double
        L.sf_batch_current_step.restype = C.c_int
        ns = (C.c_int * len(self.nstep))(*self.nstep)
        b = self.base
        self.h = C.c_long(L.sf_batch_create(b.MP, C.c_int(len(self.nstep)), ns, C.c_float(b.deltat),
                                            C.c_float(b.deltatover2), C.c_float(b.deltatsqover2)))
        self.nrec = []
        for i, (src, rec, n) in enumerate(self.shots):
            if src is not None and src.n:
                stf = np.asarray(src.stf_table(n), np.float64)
                for k in range(src.n):
                    arr = np.ascontiguousarray(src.sourcearrays[k], np.float32).reshape(-1)   # (125, 3): c fastest
                    col = np.ascontiguousarray(stf[:n, k].astype(np.float32))
                    L.sf_batch_add_source(self.h, C.c_int(i), C.c_int(int(src.ispec[k])), abi.ptr(arr), abi.ptr(col))
            nr = 0 if rec is None else rec.n
            for r in range(nr):
                hx = np.ascontiguousarray(rec.hxir[r], np.float64)
                hy = np.ascontiguousarray(rec.hetar[r], np.float64)
                hz = np.ascontiguousarray(rec.hgammar[r], np.float64)
                nu = np.ascontiguousarray(np.asarray(rec.nu[r], np.float64).ravel(order="F"))
                L.sf_batch_add_receiver(self.h, C.c_int(i), C.c_int(int(rec.ispec[r])), abi.ptr(hx), abi.ptr(hy),
                                        abi.ptr(hz), abi.ptr(nu))
            self.nrec.append(nr)
        if shots_per_group is not None or assembly is not None:
            L.sf_batch_set_options(self.h, C.c_int(shots_per_group or 4), C.c_int(2 if assembly is None else assembly))
        L.sf_batch_finalize(self.h)

    def autotune(self, trial_steps: int = 5) -> dict:
        ms = self.L.sf_batch_autotune(self.h, C.c_int(trial_steps))
        return dict(self.options(), step_ms=float(ms))

    def options(self) -> dict:
        sc, asm = C.c_int(0), C.c_int(0)
        self.L.sf_batch_get_options(self.h, C.byref(sc), C.byref(asm))
        self.L.sf_batch_get_tuning.restype = C.c_int
        return {"shots_per_group": sc.value, "assembly": asm.value,
                "lds_overlay": int(self.L.sf_batch_get_tuning(self.h, b"lds_overlay"))}

    def set_tuning(self, name: str, value: int):
        self.L.sf_batch_set_tuning(self.h, name.encode(), C.c_int(int(value)))

    def set_options(self, shots_per_group: int, assembly: int):
        self.L.sf_batch_set_options(self.h, C.c_int(shots_per_group), C.c_int(assembly))

    def run(self, nsteps: int | None = None, graph_steps: int = 10):
        n = max(self.nstep) if nsteps is None else int(nsteps)
        self.L.sf_batch_run(self.h, C.c_int(n), C.c_int(graph_steps))

    def current_step(self) -> int:
        return int(self.L.sf_batch_current_step(self.h))

    def reset(self):
        self.L.sf_batch_reset(self.h)

    def synchronize(self):
        self.base.synchronize()

    def time(self, nsteps: int) -> tuple[float, float]:
        k, s = C.c_double(0), C.c_double(0)
        self.L.sf_batch_time(self.h, C.c_int(nsteps), C.byref(k), C.byref(s))
        return k.value, s.value

    def seismograms(self, shot: int) -> np.ndarray:
        out = np.zeros((self.nstep[shot], max(self.nrec[shot], 1), 3), np.float32)
        if self.nrec[shot]:
            self.L.sf_batch_seismograms(self.h, C.c_int(shot), abi.ptr(out))
        return out[:, : self.nrec[shot]]

    def displ(self, shot: int) -> np.ndarray:
        out = np.zeros(3 * self.db.nglob, np.float32)
        self.L.sf_batch_displ(self.h, C.c_int(shot), abi.ptr(out))
        return out

    def potential(self, shot: int) -> np.ndarray:
        """Potential of one shot at every global point (zero outside the fluid elements)."""
        out = np.zeros(self.db.nglob, np.float32)
        self.L.sf_batch_potential(self.h, C.c_int(shot), abi.ptr(out))
        return out

    def close(self):
        if self.h.value:
            self.L.sf_batch_destroy(self.h)
            self.h = C.c_long(0)
        self.base.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
