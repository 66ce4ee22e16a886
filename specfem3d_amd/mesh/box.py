"""Structured box mesher + database generation for the solver loop.

Produces the same database the reference's internal mesher and
``xgenerate_databases`` produce for a box (``src/meshfem3D/``,
``src/generate_databases/``): hexahedral elements on a regular grid, GLL points
from the HEX8 map, one global number per unique point, regular-element shortcut,
Stacey faces on the chosen sides, a free surface on top, acoustic elements on top
of elastic ones with their coupling faces, and a per-GLL-point material model.

Numbering (our own; the physics does not depend on it):
  * element ``e = ex + nx * (ey + ny * ez)``, ``ez = 0`` at the bottom;
  * global point ``g = gx + NX * (gy + NY * gz)`` with ``NX = 4 nx + 1`` etc.,
    which keeps neighbouring points close in memory.
"""
from __future__ import annotations

from typing import Callable, Iterable

import numpy as np

from ..gll import get_gll
from .database import (Database, Faces, FACE_XMIN, FACE_XMAX, FACE_YMIN, FACE_YMAX,
                       FACE_ZMIN, FACE_ZMAX, classify_regular, face_geometry, jacobian_terms)

MaterialFn = Callable[[np.ndarray, np.ndarray, np.ndarray], tuple]


def homogeneous(rho: float, vp: float, vs: float) -> MaterialFn:
    def f(x, y, z):
        s = np.shape(x)
        return (np.full(s, rho, np.float32), np.full(s, vp, np.float32), np.full(s, vs, np.float32))
    return f


def build_box(nx: int, ny: int, nz: int,
              extent: tuple = (0.0, 1.0, 0.0, 1.0, -1.0, 0.0),
              material: MaterialFn | None = None,
              acoustic_top_layers: int = 0,
              acoustic_material: MaterialFn | None = None,
              stacey_faces: Iterable[int] = (FACE_XMIN, FACE_XMAX, FACE_YMIN, FACE_YMAX, FACE_ZMIN),
              free_surface: bool = True,
              warp: Callable | None = None) -> Database:
    """Build a box database.

    ``extent`` = (xmin, xmax, ymin, ymax, zmin, zmax). ``material(x, y, z)`` returns
    float32 (rho, vp, vs) arrays evaluated at GLL points (shape (nspec, 125)).
    The top ``acoustic_top_layers`` element layers use ``acoustic_material``
    (vs = 0). ``warp(x, y, z) -> (x, y, z)`` displaces anchor nodes (makes elements
    irregular, for the irregular-Jacobian path); it must keep faces planar in the
    sense that the HEX8 map stays valid.
    """
    x0, x1, y0, y1, z0, z1 = extent
    gll = get_gll()
    xi = gll.xigll
    nspec = nx * ny * nz
    NX, NY, NZ = 4 * nx + 1, 4 * ny + 1, 4 * nz + 1
    nglob = NX * NY * NZ
    if nglob >= 2 ** 31:
        raise ValueError("mesh too large for int32 global numbering")

    # global point grid coordinates along each axis (float64): element corner + GLL offset
    def axis(n, a, b):
        h = (b - a) / n
        g = np.empty(4 * n + 1)
        for e in range(n):
            base = a + e * h
            g[4 * e:4 * e + 5] = base + 0.5 * (xi + 1.0) * h
        g[-1] = b
        return g, h

    gx, hx = axis(nx, x0, x1)
    gy, hy = axis(ny, y0, y1)
    gz, hz = axis(nz, z0, z1)

    # ibool[e, k, j, i]
    ex = np.arange(nx)
    ey = np.arange(ny)
    ez = np.arange(nz)
    l = np.arange(5)
    # build per-axis global index for element e_axis and local l: 4*e + l
    IX = (4 * ex[:, None] + l[None, :])  # (nx,5)
    IY = (4 * ey[:, None] + l[None, :])
    IZ = (4 * ez[:, None] + l[None, :])
    ibool = (IX[None, None, :, None, None, :]
             + NX * (IY[None, :, None, None, :, None] + NY * IZ[:, None, None, :, None, None]))
    # axes: (ez, ey, ex, k, j, i)
    ibool = ibool.reshape(nspec, 5, 5, 5).astype(np.int32)

    # anchors (nspec, 8, 3)
    from .database import HEX8_REF
    EZ, EY, EX = np.meshgrid(ez, ey, ex, indexing="ij")
    EX = EX.ravel(); EY = EY.ravel(); EZ = EZ.ravel()
    anchors = np.empty((nspec, 8, 3))
    for a in range(8):
        rx, ry, rz = HEX8_REF[a]
        anchors[:, a, 0] = x0 + (EX + (rx + 1) / 2) * hx
        anchors[:, a, 1] = y0 + (EY + (ry + 1) / 2) * hy
        anchors[:, a, 2] = z0 + (EZ + (rz + 1) / 2) * hz

    GZ, GY, GX = np.meshgrid(gz, gy, gx, indexing="ij")
    xs = GX.ravel(); ys = GY.ravel(); zs = GZ.ravel()
    if warp is not None:
        # warped anchors; GLL points follow from the trilinear (HEX8) map of each element, which
        # agrees on shared faces, so every global point gets one consistent position
        from .database import gll_shape_tables
        ax, ay, az = warp(anchors[..., 0], anchors[..., 1], anchors[..., 2])
        anchors = np.stack([ax, ay, az], -1)
        Nsh, _ = gll_shape_tables()
        Xg = np.einsum("pa,eac->epc", Nsh, anchors)
        pos = np.zeros((nglob, 3))
        pos[ibool.reshape(nspec, 125).ravel()] = Xg.reshape(-1, 3)
        xs, ys, zs = pos[:, 0], pos[:, 1], pos[:, 2]
    xstore = xs.astype(np.float32)
    ystore = ys.astype(np.float32)
    zstore = zs.astype(np.float32)

    if warp is None:
        irregular = np.zeros(nspec, np.int32)
        cubic = np.isclose(hx, hy, rtol=1e-12) and np.isclose(hx, hz, rtol=1e-12)
        if cubic:
            nirr = 0
            h = hx
            xix_reg = np.float32(1.0 / (0.5 * h))
            jac_reg = np.float32((0.5 * h) ** 3)
        else:
            irregular, nirr, xix_reg, jac_reg = classify_regular(anchors)
    else:
        irregular, nirr, xix_reg, jac_reg = classify_regular(anchors)
    if nirr:
        irr_el = np.nonzero(irregular)[0]
        jac9, jdet = jacobian_terms(anchors[irr_el])
    else:
        jac9 = np.zeros((0, 9, 125), np.float32)
        jdet = np.zeros((0, 125), np.float32)

    # domains
    is_ac = np.zeros(nspec, bool)
    if acoustic_top_layers > 0:
        is_ac = EZ >= (nz - acoustic_top_layers)
    is_el = ~is_ac

    # materials at GLL points
    if material is None:
        material = homogeneous(2300.0, 2800.0, 1500.0)
    # the models are continuous inside each domain, so they are evaluated once per global point
    # of the domain and gathered to the elements' local points
    ib = ibool.reshape(nspec, 125)
    rho = np.empty((nspec, 125), np.float32)
    vp = np.empty((nspec, 125), np.float32)
    vs = np.empty((nspec, 125), np.float32)

    def fill(mask, fn, fluid):
        loc = ib[mask]
        seen = np.zeros(nglob, bool)
        seen[loc] = True
        pts = np.nonzero(seen)[0]                     # sorted distinct points of the domain
        lut = np.empty(nglob, np.int64)
        lut[pts] = np.arange(len(pts))
        r, p, s = fn(xstore[pts], ystore[pts], zstore[pts])
        pos = lut[loc]
        rho[mask] = np.asarray(r, np.float32)[pos]
        vp[mask] = np.asarray(p, np.float32)[pos]
        vs[mask] = 0.0 if fluid else np.asarray(s, np.float32)[pos]

    if is_el.any():
        if is_el.all():
            r, p, s = material(xstore, ystore, zstore)
            rho[:] = np.asarray(r, np.float32)[ib]
            vp[:] = np.asarray(p, np.float32)[ib]
            vs[:] = np.asarray(s, np.float32)[ib]
        else:
            fill(is_el, material, False)
    if is_ac.any():
        fill(is_ac, acoustic_material or homogeneous(1020.0, 1500.0, 0.0), True)
    f43 = np.float32(4.0 / 3.0)
    kappa = (rho * (vp * vp - f43 * vs * vs)).astype(np.float32)
    mu = (rho * vs * vs).astype(np.float32)

    db = Database(nspec=nspec, nglob=nglob, ibool=ibool, xstore=xstore, ystore=ystore, zstore=zstore,
                  irregular_element_number=irregular, nspec_irregular=nirr,
                  xix_regular=np.float32(xix_reg), jacobian_regular=np.float32(jac_reg),
                  jac9=jac9, jacobian=jdet,
                  rhostore=rho, kappastore=kappa, mustore=mu,
                  rho_vp=(rho * vp).astype(np.float32), rho_vs=(rho * vs).astype(np.float32),
                  ispec_is_elastic=is_el, ispec_is_acoustic=is_ac, anchors=anchors)
    db.meta.update(kind="box", nx=nx, ny=ny, nz=nz, extent=tuple(extent), hx=hx, hy=hy, hz=hz,
                   acoustic_top_layers=acoustic_top_layers)

    # boundary faces
    side_sel = {
        FACE_XMIN: EX == 0, FACE_XMAX: EX == nx - 1, FACE_YMIN: EY == 0,
        FACE_YMAX: EY == ny - 1, FACE_ZMIN: EZ == 0, FACE_ZMAX: EZ == nz - 1,
    }
    abs_isp, abs_if = [], []
    for f in stacey_faces:
        e = np.nonzero(side_sel[f])[0]
        abs_isp.append(e)
        abs_if.append(np.full(len(e), f))
    if abs_isp:
        isp = np.concatenate(abs_isp).astype(np.int32)
        ifc = np.concatenate(abs_if)
        ijk, nrm, jw = face_geometry(anchors, isp, ifc, xstore, ystore, zstore, ibool)
        db.abs_faces = Faces(isp, ijk, nrm, jw)
    if free_surface and FACE_ZMAX not in tuple(stacey_faces):
        e = np.nonzero(side_sel[FACE_ZMAX])[0].astype(np.int32)
        ijk, nrm, jw = face_geometry(anchors, e, np.full(len(e), FACE_ZMAX), xstore, ystore, zstore, ibool)
        db.free_faces = Faces(e, ijk, nrm, jw)
    if acoustic_top_layers > 0 and is_el.any():
        # bottom faces of the lowest acoustic layer: normal points down, away from the fluid
        e = np.nonzero(is_ac & (EZ == nz - acoustic_top_layers))[0].astype(np.int32)
        ijk, nrm, jw = face_geometry(anchors, e, np.full(len(e), FACE_ZMIN), xstore, ystore, zstore, ibool)
        db.coupling_faces = Faces(e, ijk, nrm, jw)
    return db


def gradient_grf_model(extent, vp_top=2800.0, vp_bot=5000.0, vs_top=1500.0, vs_bot=2900.0,
                       rho_top=2300.0, rho_bot=2800.0, amp=0.05, corr_len=2000.0, seed=1234,
                       nmodes=64, z_top=None) -> MaterialFn:
    """Linear-in-depth vp/vs/rho times (1 + amp * g), g a unit-variance Gaussian random field.

    The field is a seeded sum of ``nmodes`` random Fourier modes with wavenumbers
    drawn from the Gaussian spectrum of a field with correlation length
    ``corr_len`` (so g ~ N(0, 1) pointwise, correlation exp(-r^2 / (2 L^2))),
    evaluated in float64 at each GLL point; the same seed gives the same model.
    """
    x0, x1, y0, y1, z0, z1 = extent
    zt = z1 if z_top is None else z_top
    rng = np.random.default_rng(seed)
    k = rng.normal(0.0, 1.0 / corr_len, size=(nmodes, 3))
    ph = rng.uniform(0.0, 2 * np.pi, size=nmodes)
    amp_m = np.sqrt(2.0 / nmodes)

    def grf_block(x, y, z):
        out = np.zeros(np.shape(x))
        for m in range(nmodes):
            out += np.cos(k[m, 0] * x + k[m, 1] * y + k[m, 2] * z + ph[m])
        return amp_m * out

    def grf(x, y, z):
        x = np.asarray(x, np.float64); y = np.asarray(y, np.float64); z = np.asarray(z, np.float64)
        n = x.size
        if n < (1 << 20):
            return grf_block(x, y, z)
        # elementwise, so blocks on threads give the same values (numpy ufuncs release the GIL)
        import os
        from concurrent.futures import ThreadPoolExecutor
        xf, yf, zf = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        out = np.empty(n)
        step = 1 << 18
        def work(a):
            out[a:a + step] = grf_block(xf[a:a + step], yf[a:a + step], zf[a:a + step])
        with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
            list(ex.map(work, range(0, n, step)))
        return out.reshape(np.shape(x))

    def f(x, y, z):
        depth_frac = np.clip((zt - np.asarray(z, np.float64)) / max(zt - z0, 1e-30), 0.0, 1.0)
        g = 1.0 + amp * grf(x, y, z)
        vp = (vp_top + (vp_bot - vp_top) * depth_frac) * g
        vs = (vs_top + (vs_bot - vs_top) * depth_frac) * g
        rho = (rho_top + (rho_bot - rho_top) * depth_frac) * g
        return rho.astype(np.float32), vp.astype(np.float32), vs.astype(np.float32)

    f.grf = grf
    return f
