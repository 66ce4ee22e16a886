"""The solver database files of the reference (NPROC = 1 partition): ``proc000000_external_mesh.bin``
and ``proc000000_attenuation.bin``.

Format: Fortran unformatted sequential files, every ``write(IOUT) array`` one record framed by 4-byte
length markers, arrays in Fortran (column-major) order, ``CUSTOM_REAL`` float32, ``logical`` as 4-byte
integers. The record sequence is the one ``save_arrays_solver_mesh`` writes
(src/generate_databases/save_arrays_solver.F90:112-330) and ``read_mesh_databases`` reads back
(src/specfem3D/read_mesh_databases.F90:158-1100); attenuation.bin as
src/specfem3D/prepare_attenuation.f90:93-111 reads it. The C-PML records (save_arrays_solver.F90:
186-218: widths, separation distance, regions, element lists, damping profiles) are written when the
database has C-PML layers and read with ``pml=True``; the interface-point lists that only adjoint runs
with C-PML use are not. Records of other run-time switches (ocean load, anisotropy, GPU mesh
colouring) are not written and refused on reading; records without a counterpart here (poroelastic
domains, MPI interfaces) are written empty.

``write_solver_database`` also fills the bookkeeping the solver only uses outside the time loop:
boundary element lists ``ibelm_*`` from the faces' ids, the external-surface flags (faces of one
element only) and a corner-sharing element adjacency (the reference adds neighbours of neighbours;
the solver uses it to locate points). ``read_solver_database`` rebuilds a ``Database`` -- the
mass matrices are combined with their Stacey terms and inverted as prepare_timerun_mass_matrices
does -- ready for the C-ABI path.
"""
from __future__ import annotations

import numpy as np

from .database import Database, Faces, face_ijk

NGLL3 = 125


# ------------------------------------------------------------------------------ record I/O
class FortranWriter:
    def __init__(self, path):
        self.f = open(path, "wb")

    def rec(self, *arrays):
        """One record holding ``arrays`` back to back (each in Fortran order)."""
        data = b"".join(np.asarray(a).ravel(order="F").tobytes() for a in arrays)
        n = np.array([len(data)], np.int32).tobytes()
        self.f.write(n + data + n)

    def i(self, *v):
        self.rec(np.array(v, np.int32))

    def close(self):
        self.f.close()


class FortranReader:
    def __init__(self, path):
        self.buf = open(path, "rb").read()
        self.pos = 0

    def raw(self) -> bytes:
        n = int(np.frombuffer(self.buf, np.int32, 1, self.pos)[0])
        data = self.buf[self.pos + 4:self.pos + 4 + n]
        tail = int(np.frombuffer(self.buf, np.int32, 1, self.pos + 4 + n)[0])
        if tail != n:
            raise ValueError(f"corrupt Fortran record at byte {self.pos}")
        self.pos += 8 + n
        return data

    def arr(self, dtype, shape=None):
        a = np.frombuffer(self.raw(), dtype).copy()
        if shape is not None:
            a = a.reshape(shape, order="F")
        return a

    def i(self):
        return [int(v) for v in np.frombuffer(self.raw(), np.int32)]

    def f(self):
        return float(np.frombuffer(self.raw(), np.float32)[0])

    def expect(self, itest):
        v = self.i()[0]
        if v != itest:
            raise ValueError(f"database file out of sequence: expected {itest}, found {v}")


def _logical(a):
    return np.asarray(a, bool).astype(np.int32)


def _faces_fortran(f: Faces):
    """(ispec 1-based (nf), ijk (3,25,nf) 1-based, jacobian2Dw (25,nf), normal (3,25,nf)) in Fortran order."""
    return (f.ispec.astype(np.int32) + 1, np.transpose(f.ijk, (2, 1, 0)).astype(np.int32) + 1,
            np.transpose(f.jacobian2Dw, (1, 0)).astype(np.float32), np.transpose(f.normal, (2, 1, 0)).astype(np.float32))


def _faces_from_fortran(isp, ijk, jw, nrm) -> Faces:
    return Faces(np.asarray(isp, np.int32) - 1, np.ascontiguousarray(np.transpose(ijk, (2, 1, 0))).astype(np.int32) - 1,
                 np.ascontiguousarray(np.transpose(nrm, (2, 1, 0))).astype(np.float32),
                 np.ascontiguousarray(np.transpose(jw, (1, 0))).astype(np.float32))


def _face_id(ijk: np.ndarray) -> int:
    """Element face (1 xmin, 2 xmax, 3 ymin, 4 ymax, 5 bottom, 6 top) of one face's 25 (i, j, k) points."""
    for c in range(3):
        for v, fid in ((0, 2 * c + 1), (4, 2 * c + 2)):
            if np.all(ijk[:, c] == v):
                return fid
    raise ValueError("face points do not lie on one element face")


def _element_faces(ibool: np.ndarray) -> dict:
    """Corner-node key -> list of (element, face id) over the 6 faces of every element."""
    corners = {1: ((0, 0, 0), (0, 4, 0), (0, 4, 4), (0, 0, 4)), 2: ((4, 0, 0), (4, 4, 0), (4, 4, 4), (4, 0, 4)),
               3: ((0, 0, 0), (4, 0, 0), (4, 0, 4), (0, 0, 4)), 4: ((0, 4, 0), (4, 4, 0), (4, 4, 4), (0, 4, 4)),
               5: ((0, 0, 0), (4, 0, 0), (4, 4, 0), (0, 4, 0)), 6: ((0, 0, 4), (4, 0, 4), (4, 4, 4), (0, 4, 4))}
    out = {}
    for fid, cs in corners.items():
        g = np.stack([ibool[:, k, j, i] for (i, j, k) in cs], 1)
        g.sort(axis=1)
        for e, key in enumerate(map(tuple, g)):
            out.setdefault(key, []).append((e, fid))
    return out


# ------------------------------------------------------------------------------ external_mesh.bin
def write_solver_database(db: Database, path: str):
    """Write ``db`` as proc000000_external_mesh.bin (save_arrays_solver.F90:112-330); needs finalize()."""
    if not db.mass_parts and (db.nspec_elastic or db.nspec_acoustic):
        raise ValueError("finalize() the database before writing it")
    nspec, nglob, nirr = db.nspec, db.nglob, db.nspec_irregular
    acoustic, elastic = db.nspec_acoustic > 0, db.nspec_elastic > 0
    ib = np.transpose(db.ibool, (3, 2, 1, 0)).astype(np.int32) + 1          # (i, j, k, ispec)
    loc = lambda a: np.asarray(a, np.float32).reshape(-1, 5, 5, 5).transpose(3, 2, 1, 0)  # (nspec,125) -> (5,5,5,n)
    w = FortranWriter(path)
    w.i(nspec)
    w.i(nglob)
    w.i(nirr)
    w.rec(ib)
    w.rec(db.xstore.astype(np.float32))
    w.rec(db.ystore.astype(np.float32))
    w.rec(db.zstore.astype(np.float32))
    w.rec(db.irregular_element_number.astype(np.int32))
    w.rec(np.float32(db.xix_regular))
    w.rec(np.float32(db.jacobian_regular))
    jac9 = db.jac9 if nirr else np.zeros((0, 9, 125), np.float32)
    for t in range(9):
        w.rec(loc(jac9[:, t, :]))
    w.rec(loc(db.jacobian if nirr else np.zeros((0, 125), np.float32)))
    w.rec(loc(db.kappastore))
    w.rec(loc(db.mustore))
    w.rec(_logical(db.ispec_is_acoustic))
    w.rec(_logical(db.ispec_is_elastic))
    w.rec(_logical(np.zeros(nspec, bool)))
    w.i(9999)
    mp = db.mass_parts
    if acoustic:
        w.rec(mp["rmass_acoustic"])
    w.rec(loc(db.rhostore))
    if elastic:
        w.rec(mp["rmass"])
        w.rec(loc(db.rho_vp))
        w.rec(loc(db.rho_vs))
    pml = db.pml
    if pml is not None:
        w.i(pml.n)
        for c in range(3):
            w.rec(np.float32(pml.width[c]))
        w.rec(np.float32(pml.min_distance))
        if pml.n:
            w.rec(pml.regions.astype(np.int32))
            w.rec(pml.cpml_to_spec.astype(np.int32) + 1)
            w.rec(_logical(pml.is_cpml(nspec)))
            for arr in (pml.d, pml.K, pml.alpha):
                for c in range(3):
                    w.rec(loc(arr[c]))
    f = db.abs_faces
    with_faces = db.stacey or pml is not None
    w.i(f.n if with_faces else 0)
    if with_faces and f.n:
        for a in _faces_fortran(f):
            w.rec(a)
    if db.stacey and pml is None and f.n:
        if elastic:
            w.rec(mp["rmassx"]); w.rec(mp["rmassy"]); w.rec(mp["rmassz"])
        if acoustic:
            w.rec(mp["rmassz_acoustic"])
    w.i(9998)
    # boundary element lists by side (xmin, xmax, ymin, ymax, bottom, top)
    sides = {s: [] for s in range(1, 7)}
    for e, ijk in zip(f.ispec if with_faces else [], f.ijk if with_faces else []):
        sides[_face_id(ijk)].append(int(e) + 1)
    for e, ijk in zip(db.free_faces.ispec, db.free_faces.ijk):
        sides[_face_id(ijk)].append(int(e) + 1)
    for s in range(1, 7):
        w.i(len(sides[s]))
    for s in range(1, 7):
        if sides[s]:
            w.rec(np.array(sides[s], np.int32))
    for faces in (db.free_faces, db.coupling_faces):
        w.i(faces.n)
        if faces.n:
            for a in _faces_fortran(faces):
                w.rec(a)
    w.i(0)       # num_coupling_ac_po_faces
    w.i(0)       # num_coupling_el_po_faces
    w.i(9997)
    w.i(0)       # num_interfaces_ext_mesh (one partition)
    w.i(0)       # max_nibool_interfaces_ext_mesh
    w.i(9996)
    w.rec(_logical(np.ones(nspec, bool)))                       # ispec_is_inner: no MPI interfaces
    for present, flag in ((acoustic, db.ispec_is_acoustic), (elastic, db.ispec_is_elastic)):
        if present:
            el = np.nonzero(flag)[0].astype(np.int32) + 1
            w.i(len(el), 0)                                      # nspec_inner, nspec_outer
            w.i(len(el))                                         # num_phase_ispec
            ph = np.zeros((len(el), 2), np.int32)
            ph[:, 1] = el                                        # phase_ispec_inner(:, 2) = inner elements
            w.rec(ph)
    # external surface: element faces not shared with another element
    faces = _element_faces(db.ibool)
    surf_e = np.zeros(nspec, bool)
    surf_g = np.zeros(nglob, bool)
    for key, owners in faces.items():
        if len(owners) == 1:
            e, fid = owners[0]
            surf_e[e] = True
            ijk = face_ijk(fid)
            surf_g[db.ibool[e, ijk[:, 2], ijk[:, 1], ijk[:, 0]]] = True
    w.i(len(sides[6]))                                           # nfaces_surface (MOVIE_TYPE 1: top faces)
    w.rec(_logical(surf_e))
    w.rec(_logical(surf_g))
    # corner-sharing element adjacency (CSR, 1-based)
    corner = db.ibool[:, [0, 0, 0, 0, 4, 4, 4, 4], [0, 0, 4, 4, 0, 0, 4, 4], [0, 4, 0, 4, 0, 4, 0, 4]]
    by_node = {}
    for e in range(nspec):
        for g in corner[e]:
            by_node.setdefault(int(g), []).append(e)
    xadj = [0]
    adj = []
    for e in range(nspec):
        nb = sorted({o for g in corner[e] for o in by_node[int(g)]} - {e})
        adj.extend(o + 1 for o in nb)
        xadj.append(len(adj))
    w.i(len(adj))
    w.rec(np.array(xadj, np.int32))
    w.rec(np.array(adj, np.int32))
    w.i(9995)
    w.close()


def read_solver_database(path: str, dt: float, stacey: bool = True, ocean_load: bool = False, pml: bool = False,
                         anisotropy: bool = False, mesh_coloring: bool = False) -> Database:
    """Read proc000000_external_mesh.bin (read_mesh_databases.F90) into a ``Database``; the switches are
    the Par_file values that decide which records exist. HEX8 anchors are not in the file: ``anchors``
    holds each element's 8 corner GLL points, which is what the locator and face geometry use."""
    if ocean_load or anisotropy or mesh_coloring:
        raise NotImplementedError("ocean load, anisotropy and GPU mesh colouring records are not supported")
    if pml and stacey:
        raise ValueError("C-PML and Stacey conditions exclude each other")
    r = FortranReader(path)
    nspec, nglob, nirr = r.i()[0], r.i()[0], r.i()[0]
    ibool = r.arr(np.int32, (5, 5, 5, nspec)).transpose(3, 2, 1, 0) - 1           # [e, k, j, i]
    xs, ys, zs = (r.arr(np.float32) for _ in range(3))
    irregular = r.arr(np.int32)
    xix_reg, jac_reg = np.float32(r.f()), np.float32(r.f())
    per_el = lambda a, n: np.ascontiguousarray(a.reshape(5, 5, 5, n, order="F").transpose(3, 2, 1, 0).reshape(n, 125))
    terms = [per_el(r.arr(np.float32), nirr) for _ in range(9)]      # 9 records, empty without irregular elements
    jac9 = np.stack(terms, 1) if nirr else np.zeros((0, 9, 125), np.float32)
    jac = per_el(r.arr(np.float32), nirr)
    kappa, mu = per_el(r.arr(np.float32), nspec), per_el(r.arr(np.float32), nspec)
    is_ac, is_el, is_po = (r.arr(np.int32).astype(bool) for _ in range(3))
    if is_po.any():
        raise NotImplementedError("poroelastic elements are not supported")
    r.expect(9999)
    parts = {}
    if is_ac.any():
        parts["rmass_acoustic"] = r.arr(np.float32)
    rho = per_el(r.arr(np.float32), nspec)
    rho_vp = rho_vs = None
    if is_el.any():
        parts["rmass"] = r.arr(np.float32)
        rho_vp, rho_vs = per_el(r.arr(np.float32), nspec), per_el(r.arr(np.float32), nspec)
    cpml = None
    if pml:
        from ..pml import CPML
        ncp = r.i()[0]
        width = np.array([r.f() for _ in range(3)], np.float32)
        md = np.float32(r.f())
        if ncp:
            regions = r.arr(np.int32)
            cts = r.arr(np.int32) - 1
            r.arr(np.int32)                                   # is_CPML (follows from CPML_to_spec)
            stores = [per_el(r.arr(np.float32), ncp) for _ in range(9)]
        else:
            regions = cts = np.zeros(0, np.int32)
            stores = [np.zeros((0, 125), np.float32)] * 9
        cpml = CPML(cpml_to_spec=cts.astype(np.int32), regions=regions.astype(np.int32), f0=float("nan"),
                    instead_of_free_surface=False, d=np.stack(stores[0:3]), K=np.stack(stores[3:6]),
                    alpha=np.stack(stores[6:9]), width=width, min_distance=md)
    nabs = r.i()[0]
    abs_faces = Faces.empty()
    if nabs:
        abs_faces = _faces_from_fortran(r.arr(np.int32), r.arr(np.int32, (3, 25, nabs)), r.arr(np.float32, (25, nabs)),
                                        r.arr(np.float32, (3, 25, nabs)))
        if stacey and not pml:
            if is_el.any():
                parts["rmassx"], parts["rmassy"], parts["rmassz"] = (r.arr(np.float32) for _ in range(3))
            if is_ac.any():
                parts["rmassz_acoustic"] = r.arr(np.float32)
    r.expect(9998)
    n2d = [r.i()[0] for _ in range(6)]
    for n in n2d:
        if n:
            r.arr(np.int32)
    out = []
    for _ in range(2):
        nf = r.i()[0]
        out.append(_faces_from_fortran(r.arr(np.int32), r.arr(np.int32, (3, 25, nf)), r.arr(np.float32, (25, nf)),
                                       r.arr(np.float32, (3, 25, nf))) if nf else Faces.empty())
    free_faces, coupling_faces = out
    if r.i()[0] or r.i()[0]:
        raise NotImplementedError("poroelastic coupling faces are not supported")
    r.expect(9997)
    if r.i()[0] != 0:
        raise NotImplementedError("MPI interfaces present: this reader takes one partition (NPROC = 1)")
    r.i()
    r.expect(9996)
    # inner/outer and phase lists, surface flags and adjacency are not needed to step the mesh
    if rho_vp is None:
        rho_vp = np.zeros_like(rho)
        rho_vs = np.zeros_like(rho)
    corners = np.stack([xs, ys, zs], 1).astype(np.float64)[
        ibool[:, [0, 0, 0, 0, 4, 4, 4, 4], [0, 0, 4, 4, 0, 0, 4, 4], [0, 4, 4, 0, 0, 4, 4, 0]]]
    db = Database(nspec=nspec, nglob=nglob, ibool=np.ascontiguousarray(ibool), xstore=xs, ystore=ys, zstore=zs,
                  irregular_element_number=irregular, nspec_irregular=nirr, xix_regular=xix_reg,
                  jacobian_regular=jac_reg, jac9=jac9, jacobian=jac, rhostore=rho, kappastore=kappa, mustore=mu,
                  rho_vp=rho_vp, rho_vs=rho_vs, ispec_is_elastic=is_el, ispec_is_acoustic=is_ac, anchors=corners)
    db.abs_faces, db.free_faces, db.coupling_faces = abs_faces, free_faces, coupling_faces
    db.dt = float(dt)
    db.stacey = bool(stacey and nabs)
    db.mass_parts = parts
    _combine_mass(db)
    if cpml is not None:
        from ..pml import convolution_coefficients
        db.pml = convolution_coefficients(cpml, dt)
    db.meta.update(kind="database file")
    return db


def _combine_mass(db: Database):
    """prepare_timerun_mass_matrices (prepare_timerun.F90:246-345): add the Stacey terms, replace
    non-positive entries by 1 and invert, in float32."""
    p = db.mass_parts
    one = np.float32(1.0)

    def inv(m):
        m = m.copy()
        m[m <= 0] = one
        return (one / m).astype(np.float32)
    if "rmass" in p:
        st = db.stacey and "rmassx" in p
        db.rmassx, db.rmassy, db.rmassz = (inv(p["rmass"] + p[c]) if st else inv(p["rmass"])
                                           for c in ("rmassx", "rmassy", "rmassz"))
    else:
        db.rmassx = db.rmassy = db.rmassz = np.ones(db.nglob, np.float32)
    if "rmass_acoustic" in p:
        m = p["rmass_acoustic"] + p["rmassz_acoustic"] if (db.stacey and "rmassz_acoustic" in p) else p["rmass_acoustic"]
        db.rmass_acoustic = inv(m)
    else:
        db.rmass_acoustic = np.ones(db.nglob, np.float32)


# ------------------------------------------------------------------------------ attenuation.bin
def write_attenuation_bin(att, nspec: int, path: str):
    """proc000000_attenuation.bin: nspec, factor_common (N_SLS,5,5,5,nspec), scale_factor (5,5,5,nspec),
    factor_common_kappa, scale_factor_kappa (the order prepare_attenuation.f90:99-109 reads)."""
    w = FortranWriter(path)
    w.i(nspec)
    sls = lambda a: np.asarray(a, np.float32).reshape(nspec, 5, 5, 5, 3).transpose(4, 3, 2, 1, 0)
    loc = lambda a: np.asarray(a, np.float32).reshape(nspec, 5, 5, 5).transpose(3, 2, 1, 0)
    w.rec(sls(att.factor_common))
    w.rec(loc(att.scale_factor))
    w.rec(sls(att.factor_common_kappa))
    w.rec(loc(att.scale_factor_kappa))
    w.close()


def read_attenuation_bin(path: str):
    """(factor_common (nspec,125,3), scale_factor (nspec,125), factor_common_kappa, scale_factor_kappa)."""
    r = FortranReader(path)
    nspec = r.i()[0]
    sls = lambda a: np.ascontiguousarray(a.reshape(3, 5, 5, 5, nspec, order="F").transpose(4, 3, 2, 1, 0).reshape(nspec, 125, 3))
    loc = lambda a: np.ascontiguousarray(a.reshape(5, 5, 5, nspec, order="F").transpose(3, 2, 1, 0).reshape(nspec, 125))
    fc = sls(r.arr(np.float32))
    sc = loc(r.arr(np.float32))
    fck = sls(r.arr(np.float32))
    sck = loc(r.arr(np.float32))
    return fc, sc, fck, sck
