"""External (CUBIT/GEOCUBIT-format) hexahedral meshes -> solver database.

Reads the ``MESH`` directory format consumed by the reference's ``xdecompose_mesh`` /
``xgenerate_databases`` (``src/decompose_mesh/``, ``src/generate_databases/``):

* ``nodes_coords_file``: count, then ``id x y z``;
* ``mesh_file``: count, then ``ispec n1 .. n8`` (HEX8 anchors, bottom face counter-clockwise then
  top face);
* ``materials_file``: ``ispec material_id``; ``nummaterial_velocity_file``:
  ``domain_id material_id rho vp vs Q_kappa Q_mu anisotropy`` (domain 1 acoustic, 2 elastic);
* ``absorbing_surface_file_{xmin,xmax,ymin,ymax,bottom}`` and
  ``free_or_absorbing_surface_file_zmax``: count, then ``ispec n1 n2 n3 n4``.

GLL points come from the trilinear map of the anchors (float64, stored float32), global points
are merged by coordinates, and the rest (Jacobians, regular elements, faces, mass matrices)
follows the box generator's database semantics.
"""
from __future__ import annotations

import os

import numpy as np

from .database import Database, Faces, classify_regular, face_geometry, gll_shape_tables, jacobian_terms

# anchor indices (0-based, SPECFEM HEX8 order) of the 6 element faces, in face-id order 1..6
FACE_ANCHORS = {
    1: (0, 3, 7, 4),  # xmin (xi = -1)
    2: (1, 2, 6, 5),  # xmax
    3: (0, 1, 5, 4),  # ymin (eta = -1)
    4: (3, 2, 6, 7),  # ymax
    5: (0, 1, 2, 3),  # bottom (gamma = -1)
    6: (4, 5, 6, 7),  # top
}
STACEY_SIDES = ("xmin", "xmax", "ymin", "ymax", "bottom")


def _read_table(path, dtype=float):
    with open(path) as f:
        n = int(f.readline().split()[0])
        return np.loadtxt(f, dtype=dtype, max_rows=n, ndmin=2)


def read_material_props(path):
    rows = []
    for line in open(path):
        p = line.split("!")[0].split()
        if len(p) >= 7 and p[0].lstrip("-").isdigit():
            rows.append([float(v) for v in p[:7]])
    return np.array(rows)


def read_mesh_dir(mesh_dir: str) -> dict:
    """Raw arrays of a MESH directory (the same keys ``build_from_arrays`` takes)."""
    d = dict(nodes=_read_table(os.path.join(mesh_dir, "nodes_coords_file")),
             elements=_read_table(os.path.join(mesh_dir, "mesh_file"), np.int64),
             materials=np.loadtxt(os.path.join(mesh_dir, "materials_file"), dtype=np.int64, ndmin=2),
             material_props=read_material_props(os.path.join(mesh_dir, "nummaterial_velocity_file")))
    for nm in STACEY_SIDES:
        p = os.path.join(mesh_dir, f"absorbing_surface_file_{nm}")
        if os.path.exists(p):
            d["abs_" + nm] = _read_table(p, np.int64)
    p = os.path.join(mesh_dir, "free_or_absorbing_surface_file_zmax")
    if os.path.exists(p):
        d["free_zmax"] = _read_table(p, np.int64)
    p = os.path.join(mesh_dir, "absorbing_cpml_file")
    if os.path.exists(p):
        d["cpml"] = _read_table(p, np.int64)
    return d


def read_external_mesh(mesh_dir: str, dt: float, stacey: bool = True, pml: dict | None = None) -> Database:
    return build_from_arrays(read_mesh_dir(mesh_dir), dt, stacey=stacey, pml=pml)


def build_from_arrays(d, dt: float, stacey: bool = True, pml: dict | None = None) -> Database:
    """``pml`` = dict(f0=f0_FOR_PML, instead_of_free_surface=PML_INSTEAD_OF_FREE_SURFACE) turns the
    ``absorbing_cpml_file`` elements (``ispec region`` lines, region codes 1..7) into C-PML layers: the
    absorbing faces then carry the Dirichlet condition of the layers' outer edges and no Stacey term
    (``pml_set_local_dampingcoeff.f90``; C-PML and Stacey exclude each other)."""
    if pml is not None and stacey:
        raise ValueError("C-PML and Stacey conditions exclude each other")
    nodes = np.asarray(d["nodes"], np.float64)
    node_xyz = np.zeros((int(nodes[:, 0].max()) + 1, 3))
    node_xyz[nodes[:, 0].astype(np.int64)] = nodes[:, 1:4]
    elems = np.asarray(d["elements"], np.int64)
    elems = elems[np.argsort(elems[:, 0])]
    nspec = elems.shape[0]
    conn = elems[:, 1:9]
    anchors = node_xyz[conn]  # (nspec, 8, 3)

    mats = np.asarray(d["materials"], np.int64)
    mat_of = np.zeros(nspec, np.int64)
    mat_of[mats[:, 0] - 1] = mats[:, 1]

    # GLL points from the trilinear map, merged into global points by coordinates
    N, _ = gll_shape_tables()
    X = np.einsum("pa,eac->epc", N, anchors)  # (nspec, 125, 3) float64
    span = float(np.ptp(X.reshape(-1, 3), axis=0).max())
    key = np.round(X.reshape(-1, 3) / (span * 1e-9)).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.ravel()
    nglob = int(inv.max()) + 1
    # renumber global points by first appearance in element order (locality of the gathers)
    first = np.full(nglob, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(first, inv, np.arange(inv.size))
    newid = np.empty(nglob, np.int64)
    newid[np.argsort(first)] = np.arange(nglob)
    ib = newid[inv].astype(np.int32)
    xyz = np.zeros((nglob, 3))
    xyz[ib] = X.reshape(-1, 3)
    xstore, ystore, zstore = (xyz[:, c].astype(np.float32) for c in range(3))
    ibool = ib.reshape(nspec, 5, 5, 5)

    irregular, nirr, xix_reg, jac_reg = classify_regular(anchors)
    if nirr:
        jac9, jdet = jacobian_terms(anchors[irregular > 0])
    else:
        jac9 = np.zeros((0, 9, 125), np.float32)
        jdet = np.zeros((0, 125), np.float32)

    rho = np.empty((nspec, 125), np.float32)
    vp = np.empty((nspec, 125), np.float32)
    vs = np.empty((nspec, 125), np.float32)
    is_ac = np.zeros(nspec, bool)
    qmu = np.zeros((nspec, 125), np.float32)
    qk = np.zeros((nspec, 125), np.float32)
    for row in np.asarray(d["material_props"]):
        dom, mid, r, p, s, qkap, qm = row[:7]
        sel = mat_of == int(mid)
        rho[sel] = np.float32(r)
        vp[sel] = np.float32(p)
        vs[sel] = np.float32(s) if int(dom) != 1 else np.float32(0)
        qk[sel] = np.float32(qkap)
        qmu[sel] = np.float32(qm)
        if int(dom) == 1:
            is_ac[sel] = True
    f43 = np.float32(4.0 / 3.0)
    db = Database(nspec=nspec, nglob=nglob, ibool=ibool, xstore=xstore, ystore=ystore, zstore=zstore,
                  irregular_element_number=irregular, nspec_irregular=nirr, xix_regular=np.float32(xix_reg),
                  jacobian_regular=np.float32(jac_reg), jac9=jac9, jacobian=jdet, rhostore=rho,
                  kappastore=(rho * (vp * vp - f43 * vs * vs)).astype(np.float32),
                  mustore=(rho * vs * vs).astype(np.float32),
                  rho_vp=(rho * vp).astype(np.float32), rho_vs=(rho * vs).astype(np.float32),
                  ispec_is_elastic=~is_ac, ispec_is_acoustic=is_ac, anchors=anchors)
    db.qmu = qmu
    db.qkappa = qk
    db.meta.update(kind="external")

    def faces_of(rows):
        isp, ifc = [], []
        for row in np.asarray(rows, np.int64):
            e = int(row[0]) - 1
            fn = set(int(v) for v in row[1:5])
            for fid, idx in FACE_ANCHORS.items():
                if set(int(conn[e, k]) for k in idx) == fn:
                    isp.append(e)
                    ifc.append(fid)
                    break
            else:
                raise ValueError(f"face {row} not found on element {e + 1}")
        return np.array(isp, np.int32), np.array(ifc, np.int32)

    if stacey or pml is not None:
        parts = [faces_of(d["abs_" + nm]) for nm in STACEY_SIDES if ("abs_" + nm) in d]
        if pml is not None and pml.get("instead_of_free_surface") and "free_zmax" in d:
            parts.append(faces_of(d["free_zmax"]))
        if parts:
            isp = np.concatenate([p[0] for p in parts])
            ifc = np.concatenate([p[1] for p in parts])
            ijk, nrm, jw = face_geometry(anchors, isp, ifc, xstore, ystore, zstore, ibool)
            db.abs_faces = Faces(isp, ijk, nrm, jw)
    if "free_zmax" in d and not (pml is not None and pml.get("instead_of_free_surface")):
        isp, ifc = faces_of(d["free_zmax"])
        if len(isp):
            ijk, nrm, jw = face_geometry(anchors, isp, ifc, xstore, ystore, zstore, ibool)
            db.free_faces = Faces(isp, ijk, nrm, jw)
    if is_ac.any() and (~is_ac).any():
        isp, ifc = coupling_faces(conn, is_ac)
        if len(isp):
            ijk, nrm, jw = face_geometry(anchors, isp, ifc, xstore, ystore, zstore, ibool)
            db.coupling_faces = Faces(isp, ijk, nrm, jw)
    if pml is not None:
        from ..pml import set_damping
        if "cpml" not in d:
            raise ValueError("C-PML requested but the mesh has no absorbing_cpml_file")
        rows = np.asarray(d["cpml"], np.int64)
        rows = rows[np.argsort(rows[:, 0], kind="stable")]
        db.pml = set_damping(db, (rows[:, 0] - 1).astype(np.int32), rows[:, 1].astype(np.int32), pml["f0"],
                             bool(pml.get("instead_of_free_surface", False)))
    db.finalize(dt, stacey=stacey)
    return db


def coupling_faces(conn: np.ndarray, is_acoustic: np.ndarray):
    """Fluid-solid coupling faces (``get_coupling_surfaces_ac_el``, get_coupling_surfaces.f90:197-405):
    every face of an acoustic element that it shares with an elastic element, listed once on the
    acoustic element in element order (faces 1..6 per element); returns (ispec, face id). The normals
    (from ``face_geometry``) point out of the acoustic element, i.e. away from the fluid."""
    el_faces = set()
    for e in np.nonzero(~is_acoustic)[0]:
        for idx in FACE_ANCHORS.values():
            el_faces.add(tuple(sorted(int(conn[e, k]) for k in idx)))
    isp, ifc = [], []
    for e in np.nonzero(is_acoustic)[0]:
        for fid, idx in FACE_ANCHORS.items():
            if tuple(sorted(int(conn[e, k]) for k in idx)) in el_faces:
                isp.append(int(e))
                ifc.append(fid)
    return np.array(isp, np.int32), np.array(ifc, np.int32)
