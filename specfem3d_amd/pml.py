"""C-PML set-up on the host: damping profiles, recursive-convolution coefficients, PML mass.

The GPU entry point ``prepare_fields_elastic_pml_`` receives the convolution coefficients the
Fortran host precomputes; this module computes them the way the reference does, in single
precision with the reference's operation order, so the arrays handed over are the ones a
SPECFEM3D run would pass:

* ``set_damping`` restates ``pml_set_local_dampingcoeff`` (src/generate_databases/
  pml_set_local_dampingcoeff.f90:29-1936): PML widths and interface origins from the CPML
  element coordinates, the damping profile d(dist) (``pml_damping_profile_l``, :1942-1983), K = 1,
  alpha = ALPHA_MAX (1 - dist), and the separation of nearly equal parameters on edges and corners
  (:1989-2023 and the second element loop);
* ``convolution_coefficients`` restates ``prepare_timerun_pml`` (src/specfem3D/prepare_timerun.F90:
  592-1030): ``compute_convolution_coef`` (second-order recursive convolution, Xie et al. 2014),
  ``l_parameter_computation`` (pml_compute_accel_contribution.f90:445-597),
  ``lijk_parameter_computation`` and ``l{x,y,z}_parameter_computation``
  (pml_compute_memory_variables.f90:912-1283);
* ``pml_elastic_mass`` restates ``define_mass_matrices_pml_elastic``
  (src/shared/define_mass_matrices.f90:527-735).

Region codes (constants.h): 1 X, 2 Y, 3 Z, 4 XY, 5 XZ, 6 YZ, 7 XYZ. Only elastic C-PML elements
are supported on the GPU, as in the reference's GPU path (acoustic C-PML is CPU-only there).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .gll import get_gll

CPML_X_ONLY, CPML_Y_ONLY, CPML_Z_ONLY, CPML_XY_ONLY, CPML_XZ_ONLY, CPML_YZ_ONLY, CPML_XYZ = range(1, 8)
NPOWER = 1.0          # constants.h.in:233
CPML_RCOEF = 0.001    # constants.h.in:237
THETA = np.float32(0.125)  # pml_par.f90:77

_HAS = {  # directions damped in each region
    CPML_X_ONLY: (1, 0, 0), CPML_Y_ONLY: (0, 1, 0), CPML_Z_ONLY: (0, 0, 1), CPML_XY_ONLY: (1, 1, 0),
    CPML_XZ_ONLY: (1, 0, 1), CPML_YZ_ONLY: (0, 1, 1), CPML_XYZ: (1, 1, 1)}

f32 = np.float32


@dataclass
class CPML:
    """C-PML elements of a database and their damping parameters (float32, per local point)."""
    cpml_to_spec: np.ndarray          # (n,) int32, 0-based element of each C-PML element
    regions: np.ndarray               # (n,) int32 region codes 1..7
    f0: float                         # f0_FOR_PML
    instead_of_free_surface: bool     # PML_INSTEAD_OF_FREE_SURFACE
    K: np.ndarray = None              # (3, n, 125) float32: K_store_x/y/z
    d: np.ndarray = None              # (3, n, 125) float32: d_store_x/y/z
    alpha: np.ndarray = None          # (3, n, 125) float32: alpha_store_x/y/z
    width: np.ndarray = None          # (3,) float32 CPML_width_x/y/z
    origins: dict = field(default_factory=dict)
    vp_max: np.float32 = None
    min_distance: np.float32 = None   # min_distance_between_CPML_parameter
    # convolution coefficients of prepare_timerun_pml, Fortran order (coef, i, j, k, ispec_CPML)
    # stored here as (n, 125, ncoef)
    coef_alpha: np.ndarray = None     # 9: exp/convolution coefficients of alpha_x, alpha_y, alpha_z
    coef_beta: np.ndarray = None      # 9: same for beta = alpha + d / K
    coef_abar: np.ndarray = None      # 5: A_1..A_5 of the acceleration terms
    coef_strain: np.ndarray = None    # 18: A_6..A_23 of the strain terms
    dt: float = None

    @property
    def n(self) -> int:
        return int(self.cpml_to_spec.shape[0])

    def is_cpml(self, nspec: int) -> np.ndarray:
        m = np.zeros(nspec, bool)
        m[self.cpml_to_spec] = True
        return m

    def spec_to_cpml(self, nspec: int) -> np.ndarray:
        """(nspec,) int32: 1-based C-PML index of each element, 0 outside the PML (spec_to_CPML)."""
        s = np.zeros(nspec, np.int32)
        s[self.cpml_to_spec] = np.arange(1, self.n + 1, dtype=np.int32)
        return s


# ------------------------------------------------------------------------------------------
# damping profiles (pml_set_local_dampingcoeff.f90)
# ------------------------------------------------------------------------------------------
def damping_profile(dist, vp, delta):
    """pml_damping_profile_l (:1942-1983): d = -(NPOWER+1) vp log(R) / (2 delta) dist^(1.2 NPOWER),
    evaluated in double precision from the single-precision inputs, then rounded."""
    dist = np.asarray(dist, np.float32)
    if np.any(dist > f32(1.0)):
        raise ValueError("C-PML error: distance to C-PML/mesh interface is bigger than thickness of C-PML layer")
    c = ((NPOWER + 1.0) * float(vp) * math.log(CPML_RCOEF)) / (2.0 * float(delta))
    return -(c * dist.astype(np.float64) ** (1.2 * NPOWER)).astype(np.float32)


def _last_min(vals: np.ndarray, init: np.float32) -> tuple[bool, int]:
    """The Fortran scan ``if (f(x) <= f(c)) c = x`` over ``vals`` (the f values in loop order) leaves
    c at the last occurrence of the minimum, or at its initial value if no x reaches it."""
    if vals.size == 0:
        return False, -1
    m = vals.min()
    if m > init:
        return False, -1
    idx = np.nonzero(vals == m)[0]
    return True, int(idx[-1])


def set_damping(db, cpml_to_spec, regions, f0: float, instead_of_free_surface: bool = False,
                dt: float | None = None) -> CPML:
    """Damping profiles of the C-PML elements (pml_set_local_dampingcoeff.f90:29-1936)."""
    cpml_to_spec = np.asarray(cpml_to_spec, np.int32)
    regions = np.asarray(regions, np.int32)
    n = len(cpml_to_spec)
    if np.any((regions < 1) | (regions > 7)):
        raise ValueError("C-PML region codes must be 1..7")
    pml = CPML(cpml_to_spec=cpml_to_spec, regions=regions, f0=float(f0),
               instead_of_free_surface=bool(instead_of_free_surface))
    ib = db.ibool.reshape(db.nspec, 125)[cpml_to_spec]
    X = db.xstore[ib]
    Y = db.ystore[ib]
    Z = db.zstore[ib]
    pi_f0 = math.pi * float(f0)
    amax = np.array([f32(pi_f0 * float(f32(0.9))), f32(pi_f0 * 1.0), f32(pi_f0 * float(f32(1.1)))], np.float32)

    lo = [db.xstore.min(), db.ystore.min(), db.zstore.min()]
    hi = [db.xstore.max(), db.ystore.max(), db.zstore.max()]
    origin = [f32(f32(lo[a] + hi[a]) / f32(2)) for a in range(3)]
    coords = (X, Y, Z)
    # widths: scan of the C-PML points of the regions damped along each axis (:200-258)
    w_lo, w_hi = [], []
    for a in range(3):
        sel = np.array([_HAS[int(r)][a] for r in regions], bool)
        c = coords[a][sel].ravel()
        o = origin[a]
        dif = (c - o).astype(np.float32)
        right = dif > 0
        c_hi = f32(hi[a])
        ok, i = _last_min(dif[right], f32(c_hi - o))
        if ok:
            c_hi = c[right][i]
        c_lo = f32(lo[a])
        ok, i = _last_min(np.abs(dif[~right]), abs(f32(c_lo - o)))
        if ok:
            c_lo = c[~right][i]
        w_hi.append(f32(f32(hi[a]) - c_hi))
        w_lo.append(f32(c_lo - f32(lo[a])))
    origin_lo = [f32(f32(lo[a]) + w_lo[a]) for a in range(3)]
    origin_hi = [f32(f32(hi[a]) - w_hi[a]) for a in range(3)]
    width = np.array([max(w_lo[a], w_hi[a]) for a in range(3)], np.float32)
    pml.width = width
    pml.origins = dict(x_left=origin_lo[0], x_right=origin_hi[0], y_back=origin_lo[1], y_front=origin_hi[1],
                       z_bottom=origin_lo[2], z_top=origin_hi[2] if instead_of_free_surface else None,
                       origin=tuple(origin))
    # maximum P velocity in the PML (used everywhere in the profile)
    vp = (db.rho_vp[cpml_to_spec] / db.rhostore[cpml_to_spec]).astype(np.float32)
    vp_max = f32(vp.max()) if n else f32(0)
    pml.vp_max = vp_max

    K = np.ones((3, n, 125), np.float32)
    d = np.zeros((3, n, 125), np.float32)
    al = np.zeros((3, n, 125), np.float32)
    for a in range(3):
        act = np.array([_HAS[int(r)][a] for r in regions], bool)
        if not act.any():
            continue
        c = coords[a][act]
        dif = (c - origin[a]).astype(np.float32)
        if np.any(dif == 0):
            raise ValueError("there is error in mesh of CPML-layer (point on the domain's mid-plane)")
        top = dif > 0
        absc = np.where(top, (c - origin_hi[a]).astype(np.float32), (origin_lo[a] - c).astype(np.float32))
        if a == 2 and not instead_of_free_surface and top.any():
            # the reference leaves the top-side profile undefined without PML_INSTEAD_OF_FREE_SURFACE
            raise ValueError("C-PML element damped along z above the mid-plane needs PML_INSTEAD_OF_FREE_SURFACE")
        dist = (absc / width[a]).astype(np.float32)
        dd = damping_profile(dist, vp_max, width[a])
        aa = (float(amax[a]) * (1.0 - dist.astype(np.float64))).astype(np.float32)
        neg = dd < 0
        dd[neg] = 0
        if np.any(aa < 0):
            raise ValueError("there is error in mesh of CPML-layer (negative alpha)")
        d[a][act] = dd
        al[a][act] = aa

    # separation of nearly equal parameters (:1100-1520): minimum GLL spacing in the PML
    if n:
        P = np.stack([X, Y, Z], -1).reshape(n, 5, 5, 5, 3)  # [e, k, j, i, c]
        base = P[:, :4, :4, :4]
        dmin = np.float32(np.inf)
        for sl in (P[:, :4, :4, 1:5], P[:, :4, 1:5, :4], P[:, 1:5, :4, :4]):
            diff = (base - sl).astype(np.float32)
            s = (diff[..., 0] * diff[..., 0] + diff[..., 1] * diff[..., 1]).astype(np.float32)
            s = (s + diff[..., 2] * diff[..., 2]).astype(np.float32)
            dmin = min(dmin, s.min())
        dmin = f32(np.sqrt(f32(dmin)))
        if dmin <= 0:
            raise ValueError("error: GLL points minimum distance")
        md = f32(f32(f32(amax[0] * dmin) / f32(width.max())) / f32(8))
    else:
        md = f32(0)
    pml.min_distance = md
    _separate(regions, K, d, al, md)
    pml.K, pml.d, pml.alpha = K, d, al
    if dt is not None:
        convolution_coefficients(pml, dt)
    return pml


def _separate(regions, K, d, al, md):
    """Second loop of pml_set_local_dampingcoeff: on edge and corner regions, parameters closer than
    ``md`` are pushed apart by 2 or 4 md, then d = (beta - alpha) K is recomputed."""
    c2 = f32(md * f32(2))
    c4 = f32(md * f32(4))
    pairs = {CPML_XY_ONLY: (0, 1), CPML_XZ_ONLY: (0, 2), CPML_YZ_ONLY: (1, 2)}
    for reg, (a, b) in pairs.items():
        sel = regions == reg
        if not sel.any():
            continue
        ka, da, aa = K[a][sel], d[a][sel], al[a][sel]
        kb, db_, ab = K[b][sel], d[b][sel], al[b][sel]
        flat = [v.ravel().copy() for v in (ka, da, aa, kb, db_, ab)]
        ka, da, aa, kb, db_, ab = flat
        for q in range(aa.size):
            A, B = aa[q], ab[q]
            if abs(f32(A - B)) < md:
                if A >= B:
                    A = f32(B + c2)
                else:
                    B = f32(A + c2)
            bA = f32(A + f32(da[q] / ka[q]))
            bB = f32(B + f32(db_[q] / kb[q]))
            if abs(f32(bA - B)) < md:
                bA = f32(B + c2) if bA >= B else f32(B + c4)
            if abs(f32(bB - A)) < md:
                bB = f32(A + c2) if bB >= A else f32(A + c4)
            if abs(f32(A - B)) < md or abs(f32(bA - B)) < md or abs(f32(bB - A)) < md:
                raise ValueError("there is an error in the separation of C-PML parameters")
            da[q] = f32(f32(bA - A) * ka[q])
            db_[q] = f32(f32(bB - B) * kb[q])
            aa[q] = A
            ab[q] = B
        shape = K[a][sel].shape
        d[a][sel] = da.reshape(shape)
        al[a][sel] = aa.reshape(shape)
        d[b][sel] = db_.reshape(shape)
        al[b][sel] = ab.reshape(shape)
    sel = regions == CPML_XYZ
    if sel.any():
        shape = K[0][sel].shape
        vals = [[v[sel].ravel().copy() for v in (K[a], d[a], al[a])] for a in range(3)]
        for q in range(vals[0][0].size):
            k = [vals[a][0][q] for a in range(3)]
            dd = [vals[a][1][q] for a in range(3)]
            a_ = [vals[a][2][q] for a in range(3)]
            a_, bt = _separate_xyz(k, dd, a_, md, c2, c4)
            for a in range(3):
                vals[a][1][q] = f32(f32(bt[a] - a_[a]) * k[a])
                vals[a][2][q] = a_[a]
        for a in range(3):
            d[a][sel] = vals[a][1].reshape(shape)
            al[a][sel] = vals[a][2].reshape(shape)


def _separate_xyz(k, d, a, md, c2, c4):
    """Corner separation (XYZ region): alpha pairs (xy, xz, yz in turn), then each beta against the
    other two alphas."""
    a = list(a)

    def pair(i, j, o):
        # alpha_i, alpha_j too close: separate them, then place alpha_o relative to the pair
        if a[i] > a[j]:
            a[i] = f32(a[j] + c2)
        else:
            a[j] = f32(a[i] + c2)
        mx, mn = max(a[i], a[j]), min(a[i], a[j])
        if a[o] > mx:
            if abs(f32(a[o] - mx)) < md:
                a[o] = f32(mx + c2)
        elif a[o] < mn:
            if abs(f32(a[o] - mn)) < md:
                if a[i] > a[j]:
                    a[i] = f32(a[o] + c4)
                    a[j] = f32(a[o] + c2)
                else:
                    a[j] = f32(a[o] + c4)
                    a[i] = f32(a[o] + c2)
        else:
            if a[i] > a[j]:
                a[i] = f32(a[j] + c4)
                a[o] = f32(a[j] + c2)
            else:
                a[j] = f32(a[i] + c4)
                a[o] = f32(a[i] + c2)

    if abs(f32(a[0] - a[1])) < md:
        pair(0, 1, 2)
    if abs(f32(a[0] - a[2])) < md:
        pair(0, 2, 1)
    if abs(f32(a[1] - a[2])) < md:
        pair(1, 2, 0)
    if abs(f32(a[0] - a[1])) < md or abs(f32(a[1] - a[2])) < md or abs(f32(a[0] - a[2])) < md:
        raise ValueError("error in separation of alpha_x, alpha_y, alpha_z")
    bt = []
    for i, (j, l) in enumerate(((1, 2), (0, 2), (0, 1))):
        b = f32(a[i] + f32(d[i] / k[i]))
        mx, mn = max(a[j], a[l]), min(a[j], a[l])
        if b > mx:
            if abs(f32(b - mx)) < md:
                b = f32(mx + c2)
        elif b < mn:
            if abs(f32(b - mn)) < md:
                b = f32(a[j] + c2) if a[j] > a[l] else f32(a[l] + c2)
        else:
            if abs(f32(b - mx)) < md:
                b = f32(mx + c2)
            if abs(f32(b - mn)) < md:
                b = f32(mn + c2)
                if abs(f32(b - mx)) < md:
                    b = f32(mx + c2)
        bt.append(b)
    for i, (j, l) in enumerate(((1, 2), (0, 2), (0, 1))):
        if abs(f32(bt[i] - a[j])) < md or abs(f32(bt[i] - a[l])) < md:
            raise ValueError("there is an error in the separation of beta and alpha")
    return a, bt


# ------------------------------------------------------------------------------------------
# convolution coefficients (prepare_timerun.F90:592-1030 and the l*_parameter_computation routines)
# ------------------------------------------------------------------------------------------
def _expf(x: np.ndarray) -> np.ndarray:
    """single-precision exp: the double-precision exponential of the float argument, rounded"""
    return np.exp(np.asarray(x, np.float32).astype(np.float64)).astype(np.float32)


def convolution_coef(bb: np.ndarray, deltat: np.float32, md: np.float32):
    """compute_convolution_coef (prepare_timerun.F90:926-1030), second-order scheme: coef0 =
    exp(-b dt), coef1 = (1 - exp(-b dt/2)) / b, coef2 = coef1 exp(-b dt/2); third-order Taylor
    expansions below |b| = md."""
    bb = np.asarray(bb, np.float32)
    dt = f32(deltat)
    temp = _expf((f32(-0.5) * bb).astype(np.float32) * dt)
    coef0 = (temp * temp).astype(np.float32)
    big = np.abs(bb) >= md
    with np.errstate(divide="ignore", invalid="ignore"):
        c1b = ((f32(1) - temp) / bb).astype(np.float32)
    c2b = (c1b * temp).astype(np.float32)
    dt2 = f32(dt * dt)
    dt3 = f32(dt2 * dt)
    dt4 = f32(dt2 * dt2)
    dth = f32(dt * f32(0.5))
    o8, o48, o384 = f32(0.125), f32(1.0) / f32(48.0), f32(1.0) / f32(384.0)
    fa, fb, fc = f32(3.0) / f32(8.0), f32(7.0) / f32(48.0), f32(5.0) / f32(128.0)
    b2 = (bb * bb).astype(np.float32)
    b3 = (b2 * bb).astype(np.float32)
    t1 = (-(f32(o8 * dt2) * bb) + f32(o48 * dt3) * b2).astype(np.float32)
    c1s = (dth + (t1 - f32(o384 * dt4) * b3).astype(np.float32)).astype(np.float32)
    t2 = (-(f32(fa * dt2) * bb) + f32(fb * dt3) * b2).astype(np.float32)
    c2s = (dth + (t2 - f32(fc * dt4) * b3).astype(np.float32)).astype(np.float32)
    coef1 = np.where(big, c1b, c1s).astype(np.float32)
    coef2 = np.where(big, c2b, c2s).astype(np.float32)
    return coef0, coef1, coef2


def _ok(cond) -> bool:
    return bool(np.all(cond))


def _l_parameters(region, kx, dx, ax, ky, dy, ay, kz, dz, az, md):
    """l_parameter_computation (pml_compute_accel_contribution.f90:445-597): A_1..A_5 for the points
    of one region (float32 arrays, one operation at a time)."""
    bx = f32(ax + f32(dx / kx))
    by = f32(ay + f32(dy / ky))
    bz = f32(az + f32(dz / kz))
    z = np.zeros_like(ax)
    if region == CPML_XYZ:
        if not _ok((abs(f32(ax - ay)) >= md) & (abs(f32(ax - az)) >= md) & (abs(f32(ay - az)) >= md)):
            raise ValueError("Error occured in l_parameter_computation in CPML_XYZ region")
        A0 = f32(f32(kx * ky) * kz)
        A1 = f32(A0 * f32(f32(f32(f32(f32(bx + by) + bz) - ax) - ay) - az))
        A2 = f32(f32(f32(A0 * f32(bx - ax)) * f32(f32(by - ay) - ax))
                 + f32(f32(A0 * f32(by - ay)) * f32(f32(bz - az) - ay)))
        A2 = f32(A2 + f32(f32(A0 * f32(bz - az)) * f32(f32(bx - ax) - az)))
        A3 = f32(f32(f32(f32(f32(A0 * f32(ax * ax)) * f32(bx - ax)) * f32(by - ax)) * f32(bz - ax))
                 / f32(f32(ay - ax) * f32(az - ax)))
        A4 = f32(f32(f32(f32(f32(A0 * f32(ay * ay)) * f32(bx - ay)) * f32(by - ay)) * f32(bz - ay))
                 / f32(f32(ax - ay) * f32(az - ay)))
        A5 = f32(f32(f32(f32(f32(A0 * f32(az * az)) * f32(bx - az)) * f32(by - az)) * f32(bz - az))
                 / f32(f32(ay - az) * f32(ax - az)))
    elif region == CPML_XY_ONLY:
        if not _ok(abs(f32(ax - ay)) >= md):
            raise ValueError("Error occured in l_parameter_computation in CPML_XY_ONLY region")
        A0 = f32(kx * ky)
        A1 = f32(A0 * f32(f32(f32(bx + by) - ax) - ay))
        A2 = f32(f32(f32(A0 * f32(bx - ax)) * f32(f32(by - ay) - ax)) - f32(f32(A0 * f32(by - ay)) * ay))
        A3 = f32(f32(f32(f32(A0 * f32(ax * ax)) * f32(bx - ax)) * f32(by - ax)) / f32(ay - ax))
        A4 = f32(f32(f32(f32(A0 * f32(ay * ay)) * f32(bx - ay)) * f32(by - ay)) / f32(ax - ay))
        A5 = z
    elif region == CPML_XZ_ONLY:
        if not _ok(abs(f32(ax - az)) >= md):
            raise ValueError("Error occured in l_parameter_computation in CPML_XZ_ONLY region")
        A0 = f32(kx * kz)
        A1 = f32(A0 * f32(f32(f32(bx + bz) - ax) - az))
        A2 = f32(f32(f32(A0 * f32(bx - ax)) * f32(-ax)) + f32(f32(A0 * f32(bz - az)) * f32(f32(bx - ax) - az)))
        A3 = f32(f32(f32(f32(A0 * f32(ax * ax)) * f32(bx - ax)) * f32(bz - ax)) / f32(az - ax))
        A4 = z
        A5 = f32(f32(f32(f32(A0 * f32(az * az)) * f32(bx - az)) * f32(bz - az)) / f32(ax - az))
    elif region == CPML_YZ_ONLY:
        if not _ok(abs(f32(ay - az)) >= md):
            raise ValueError("Error occured in l_parameter_computation in CPML_YZ_ONLY region")
        A0 = f32(ky * kz)
        A1 = f32(A0 * f32(f32(f32(by + bz) - ay) - az))
        A2 = f32(f32(f32(A0 * f32(by - ay)) * f32(f32(bz - az) - ay)) - f32(f32(A0 * f32(bz - az)) * az))
        A3 = z
        A4 = f32(f32(f32(f32(A0 * f32(ay * ay)) * f32(by - ay)) * f32(bz - ay)) / f32(az - ay))
        A5 = f32(f32(f32(f32(A0 * f32(az * az)) * f32(by - az)) * f32(bz - az)) / f32(ay - az))
    elif region == CPML_X_ONLY:
        A0 = kx
        A1 = f32(A0 * f32(bx - ax))
        A2 = f32(-f32(f32(A0 * ax) * f32(bx - ax)))
        A3 = f32(f32(A0 * f32(ax * ax)) * f32(bx - ax))
        A4, A5 = z, z
    elif region == CPML_Y_ONLY:
        A0 = ky
        A1 = f32(A0 * f32(by - ay))
        A2 = f32(-f32(f32(A0 * ay) * f32(by - ay)))
        A4 = f32(f32(A0 * f32(ay * ay)) * f32(by - ay))
        A3, A5 = z, z
    else:
        A0 = kz
        A1 = f32(A0 * f32(bz - az))
        A2 = f32(-f32(f32(A0 * az) * f32(bz - az)))
        A5 = f32(f32(A0 * f32(az * az)) * f32(bz - az))
        A3, A4 = z, z
    return A1, A2, A3, A4, A5


# region codes seen through the argument permutation of lijk_parameter_computation
_LIJK_MAP = {123: (1, 2, 3, 4, 5, 6), 132: (1, 3, 2, 5, 4, 6), 231: (3, 2, 1, 6, 5, 4)}


def _lijk(kx, dx, ax, ky, dy, ay, kz, dz, az, region, index, md):
    """lijk_parameter_computation (pml_compute_memory_variables.f90:912-1070): A_0..A_3."""
    X, Y, Zt, XY, XZ, YZ = _LIJK_MAP[index]
    bx = f32(ax + f32(dx / kx))
    by = f32(ay + f32(dy / ky))
    bz = f32(az + f32(dz / kz))
    z = np.zeros_like(ax)
    if region == CPML_XYZ:
        if not _ok((abs(f32(ax - ay)) >= md) & (abs(f32(ax - bz)) >= md) & (abs(f32(ay - bz)) >= md)):
            raise ValueError("Error in lijk_parameter_computation XYZ")
        A0 = f32(f32(kx * ky) / kz)
        A1 = f32(-f32(f32(f32(A0 * f32(ax - az)) * f32(ax - bx)) * f32(ax - by)) / f32(f32(ax - ay) * f32(ax - bz)))
        A2 = f32(-f32(f32(f32(A0 * f32(ay - az)) * f32(ay - bx)) * f32(ay - by)) / f32(f32(ay - ax) * f32(ay - bz)))
        A3 = f32(-f32(f32(f32(A0 * f32(bz - az)) * f32(bz - bx)) * f32(bz - by)) / f32(f32(bz - ax) * f32(bz - ay)))
    elif region == YZ:
        if not _ok(abs(f32(ay - bz)) >= md):
            raise ValueError("Error in lijk_parameter_computation YZ")
        A0 = f32(ky / kz)
        A1 = z
        A2 = f32(-f32(f32(A0 * f32(ay - az)) * f32(ay - by)) / f32(ay - bz))
        A3 = f32(-f32(f32(A0 * f32(bz - az)) * f32(bz - by)) / f32(bz - ay))
    elif region == XZ:
        if not _ok(abs(f32(ax - bz)) >= md):
            raise ValueError("Error in lijk_parameter_computation XZ")
        A0 = f32(kx / kz)
        A1 = f32(-f32(f32(A0 * f32(ax - az)) * f32(ax - bx)) / f32(ax - bz))
        A2 = z
        A3 = f32(-f32(f32(A0 * f32(bz - az)) * f32(bz - bx)) / f32(bz - ax))
    elif region == XY:
        if not _ok(abs(f32(ax - ay)) >= md):
            raise ValueError("Error in lijk_parameter_computation XY")
        A0 = f32(kx * ky)
        A1 = f32(-f32(f32(A0 * f32(ax - bx)) * f32(ax - by)) / f32(ax - ay))
        A2 = f32(-f32(f32(A0 * f32(ay - bx)) * f32(ay - by)) / f32(ay - ax))
        A3 = z
    elif region == X:
        A0 = kx
        A1 = f32(-f32(A0 * f32(ax - bx)))
        A2, A3 = z, z
    elif region == Y:
        A0 = ky
        A2 = f32(-f32(A0 * f32(ay - by)))
        A1, A3 = z, z
    else:
        A0 = f32(f32(1) / kz)
        A3 = f32(-f32(A0 * f32(bz - az)))
        A1, A2 = z, z
    return A0, A1, A2, A3


def _l_axis(k, d, a, region, axis):
    """l{x,y,z}_parameter_computation (pml_compute_memory_variables.f90:1076-1283): A_0, A_1."""
    if _HAS[int(region)][axis]:
        b = f32(a + f32(d / k))
        return k, f32(-f32(k * f32(a - b)))
    return np.ones_like(a), np.zeros_like(a)


def convolution_coefficients(pml: CPML, dt: float) -> CPML:
    """pml_convolution_coef_alpha/beta/abar/strain of prepare_timerun_pml (prepare_timerun.F90:782-900)."""
    n = pml.n
    deltat = f32(dt)
    md = pml.min_distance
    K, d, al = pml.K, pml.d, pml.alpha
    ca = np.zeros((n, 125, 9), np.float32)
    cb = np.zeros((n, 125, 9), np.float32)
    for a in range(3):
        c0, c1, c2 = convolution_coef(al[a], deltat, md)
        ca[..., 3 * a], ca[..., 3 * a + 1], ca[..., 3 * a + 2] = c0, c1, c2
        beta = (al[a] + (d[a] / K[a]).astype(np.float32)).astype(np.float32)
        c0, c1, c2 = convolution_coef(beta, deltat, md)
        cb[..., 3 * a], cb[..., 3 * a + 1], cb[..., 3 * a + 2] = c0, c1, c2
    cab = np.zeros((n, 125, 5), np.float32)
    cst = np.zeros((n, 125, 18), np.float32)
    for reg in np.unique(pml.regions):
        sel = pml.regions == reg
        reg = int(reg)
        kx, ky, kz = K[0][sel], K[1][sel], K[2][sel]
        dx, dy, dz = d[0][sel], d[1][sel], d[2][sel]
        ax, ay, az = al[0][sel], al[1][sel], al[2][sel]
        cab[sel] = np.stack(_l_parameters(reg, kx, dx, ax, ky, dy, ay, kz, dz, az, md), -1)
        cst[sel, :, 0:4] = np.stack(_lijk(kz, dz, az, ky, dy, ay, kx, dx, ax, reg, 231, md), -1)
        cst[sel, :, 4:8] = np.stack(_lijk(kx, dx, ax, kz, dz, az, ky, dy, ay, reg, 132, md), -1)
        cst[sel, :, 8:12] = np.stack(_lijk(kx, dx, ax, ky, dy, ay, kz, dz, az, reg, 123, md), -1)
        cst[sel, :, 12:14] = np.stack(_l_axis(kx, dx, ax, reg, 0), -1)
        cst[sel, :, 14:16] = np.stack(_l_axis(ky, dy, ay, reg, 1), -1)
        cst[sel, :, 16:18] = np.stack(_l_axis(kz, dz, az, reg, 2), -1)
    pml.coef_alpha, pml.coef_beta, pml.coef_abar, pml.coef_strain = ca, cb, cab, cst
    pml.dt = float(dt)
    return pml


# ------------------------------------------------------------------------------------------
# mass matrix with the C-PML term (define_mass_matrices.f90:527-735)
# ------------------------------------------------------------------------------------------
def pml_elastic_mass(db, pml: CPML, dt: float) -> np.ndarray:
    """Elastic mass (not inverted): non-PML elastic elements in element order, then the C-PML
    elements in C-PML order with the factor K + d dt/2 (edges: Kx Ky + (dx Ky + dy Kx) dt/2, ...)."""
    from .mesh.database import _ordered_scatter_add
    gll = get_gll()
    w = gll.wgll
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()
    nspec = db.nspec
    ib = db.ibool.reshape(nspec, 125)
    jac = np.empty((nspec, 125), np.float32)
    reg = db.irregular_element_number == 0
    jac[reg] = db.jacobian_regular
    if db.nspec_irregular:
        jac[~reg] = db.jacobian[db.irregular_element_number[~reg] - 1]
    is_pml = pml.is_cpml(nspec)
    el = db.ispec_is_elastic & ~is_pml
    base = jac.astype(np.float64) * w3[None, :] * db.rhostore.astype(np.float64)
    t_el = base[el].astype(np.float32)
    deltat = float(f32(dt))
    ce = pml.cpml_to_spec
    K = pml.K.astype(np.float64)
    d = pml.d.astype(np.float64)
    fac = np.empty((pml.n, 125))
    for q, r in enumerate(pml.regions):
        r = int(r)
        kx, ky, kz = K[0, q], K[1, q], K[2, q]
        dx, dy, dz = d[0, q], d[1, q], d[2, q]
        if r == CPML_X_ONLY:
            fac[q] = kx + dx * deltat / 2.0
        elif r == CPML_Y_ONLY:
            fac[q] = ky + dy * deltat / 2.0
        elif r == CPML_Z_ONLY:
            fac[q] = kz + dz * deltat / 2.0
        elif r == CPML_XY_ONLY:
            fac[q] = kx * ky + (dx * ky + dy * kx) * deltat / 2.0
        elif r == CPML_XZ_ONLY:
            fac[q] = kx * kz + (dx * kz + dz * kx) * deltat / 2.0
        elif r == CPML_YZ_ONLY:
            fac[q] = ky * kz + (dy * kz + dz * ky) * deltat / 2.0
        else:
            fac[q] = kx * ky * kz + (dx * ky * kz + dy * kx * kz + dz * kx * ky) * deltat / 2.0
    el_pml = db.ispec_is_elastic[ce]
    t_pml = (base[ce] * fac)[el_pml].astype(np.float32)
    idx = np.concatenate([ib[el].ravel(), ib[ce][el_pml].ravel()])
    vals = np.concatenate([t_el.ravel(), t_pml.ravel()])
    return _ordered_scatter_add(db.nglob, idx, vals)


# ------------------------------------------------------------------------------------------
# C-PML layers of a structured box (the regions utils/CPML's converter assigns)
# ------------------------------------------------------------------------------------------
def box_regions(db, layers: int, faces=("xmin", "xmax", "ymin", "ymax", "bottom")):
    """(cpml_to_spec, regions) for the outer ``layers`` element layers of a box database built by
    ``mesh.box.build_box`` on the listed sides (an element is damped along each axis on which it
    lies in a layer)."""
    nx, ny, nz = db.meta["nx"], db.meta["ny"], db.meta["nz"]
    e = np.arange(db.nspec)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    inx = (("xmin" in faces) & (ex < layers)) | (("xmax" in faces) & (ex >= nx - layers))
    iny = (("ymin" in faces) & (ey < layers)) | (("ymax" in faces) & (ey >= ny - layers))
    inz = (("bottom" in faces) & (ez < layers)) | (("top" in faces) & (ez >= nz - layers))
    code = np.zeros(db.nspec, np.int32)
    table = {(1, 0, 0): 1, (0, 1, 0): 2, (0, 0, 1): 3, (1, 1, 0): 4, (1, 0, 1): 5, (0, 1, 1): 6, (1, 1, 1): 7}
    for key, v in table.items():
        code[(inx == bool(key[0])) & (iny == bool(key[1])) & (inz == bool(key[2]))] = v
    sel = np.nonzero(code)[0].astype(np.int32)
    return sel, code[sel]
