"""Sources, receivers and source-time functions (host side of the solver loop).

Restates the reference's host setup that produces the arrays crossing the C-ABI
(``prepare_constants_device``: ``sourcearrays``, ``ispec_selected_source``,
``ispec_selected_rec``, ``hxir/hetar/hgammar``, ``nu_rec``/``nu_source``) and the
per-step source-time-function values (``stf_pre_compute`` of
``compute_add_sources_el_cuda``):

* CMTSOLUTION reader ``src/specfem3D/get_cmt.f90:28-300`` (moment tensor x 1e-7,
  ``hdur >= 5 DT``, single-source ``tshift = 0``);
  Cartesian mapping Mrr=Mzz, Mtt=Myy, Mpp=Mxx, Mrt=-Myz, Mrp=Mxz, Mtp=-Mxy
  (``locate_source.F90:210-223``).
* point location ``locate_point.f90:32-624`` (nearest element, nearest GLL point,
  Newton iterations on the HEX8 map with the 1/3 step damping and +-1.01 clamp,
  adjacent-element retry).
* ``compute_arrays_source_cmt`` ``compute_arrays_source.f90:29-149`` (interpolated
  inverse Jacobian, ``M . grad(h)`` per GLL point, float64 cast to float32).
* source time functions ``comp_source_time_function.f90`` (error-function
  Heaviside with ``hdur_Gaussian = hdur / 1.628``, Ricker with ``hdur`` = f0) and
  t0 = -2 min(tshift - hdur) or -1.2 min(tshift - 1/f0)
  (``setup_sources_receivers.f90:439-491``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import erf

from .gll import get_gll, lagrange_any
from .mesh.database import Database, hex8_shape

SOURCE_DECAY_MIMIC_TRIANGLE = 1.628


# ----------------------------------------------------------------------------- STFs
def stf_heaviside(t, hdur_gaussian):
    """Error-function step ``0.5 (1 + erf(t / hdur))`` (``comp_source_time_function``)."""
    return 0.5 * (1.0 + erf(np.asarray(t, np.float64) / hdur_gaussian))


def stf_ricker(t, f0):
    """Ricker ``(1 - 2 a t^2) exp(-a t^2)``, ``a = pi^2 f0^2`` (``comp_source_time_function_rickr``)."""
    t = np.asarray(t, np.float64)
    a = math.pi ** 2 * f0 ** 2
    return (1.0 - 2.0 * a * t * t) * np.exp(-a * t * t)


def stf_gaussian(t, hdur):
    """Gaussian ``exp(-t^2/h^2) / (sqrt(pi) h)`` (``comp_source_time_function_gauss``)."""
    t = np.asarray(t, np.float64)
    a = 1.0 / hdur ** 2
    return np.exp(-a * t ** 2) / (math.sqrt(math.pi) * hdur)


def stf_d2gaussian(t, hdur):
    """Second time derivative of ``stf_gaussian``: ``2 a (2 a t^2 - 1) exp(-a t^2) / (sqrt(pi) h)``,
    ``a = 1/h^2`` (``comp_source_time_function_d2gau``)."""
    t = np.asarray(t, np.float64)
    a = 1.0 / hdur ** 2
    return 2.0 * a * (-1.0 + 2.0 * a * t ** 2) * np.exp(-a * t ** 2) / (math.sqrt(math.pi) * hdur)


def stf_d2ricker(t, f0):
    """Second time derivative of ``stf_ricker``: ``-2 a (3 - 12 a t^2 + 4 a^2 t^4) exp(-a t^2)``
    (``comp_source_time_function_d2rck``)."""
    t = np.asarray(t, np.float64)
    a = math.pi ** 2 * np.asarray(f0, np.float64) ** 2
    return -2.0 * a * (3.0 - 12.0 * a * t * t + 4.0 * a ** 2 * t * t * t * t) * np.exp(-a * t * t)


_STF = {"ricker": lambda t, h: stf_ricker(t, h),
        "d2ricker": lambda t, h: stf_d2ricker(t, h),
        "heaviside": lambda t, h: stf_heaviside(t, h / SOURCE_DECAY_MIMIC_TRIANGLE),
        "gaussian": lambda t, h: stf_gaussian(t, h / SOURCE_DECAY_MIMIC_TRIANGLE),
        "d2gaussian": lambda t, h: stf_d2gaussian(t, h / SOURCE_DECAY_MIMIC_TRIANGLE)}


# ----------------------------------------------------------------------------- inputs
@dataclass
class CMTSource:
    x: float                    # longorUTM
    y: float                    # latorUTM
    depth_km: float
    tshift: float
    hdur: float                 # half duration (or f0 for Ricker)
    # Cartesian moment tensor (N m), already x 1e-7 and mapped from r/t/p
    Mxx: float
    Myy: float
    Mzz: float
    Mxy: float
    Mxz: float
    Myz: float
    name: str = ""


def _lines(path_or_file):
    if hasattr(path_or_file, "read"):
        return path_or_file.read().splitlines()
    with open(path_or_file) as f:
        return f.read().splitlines()


def read_cmtsolution(path, DT: float) -> list[CMTSource]:
    """Parse a CMTSOLUTION file (path or file object; one or more 13-line blocks)."""
    lines = _lines(path)
    blocks = []
    i = 0
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        blk = lines[i:i + 13]
        i += 13
        vals = {}
        for l in blk[1:]:
            if ":" in l:
                k, v = l.split(":", 1)
                vals[k.strip().lower()] = v.strip()
        blocks.append(vals)
    out = []
    for v in blocks:
        mrr, mtt, mpp = (float(v[k]) for k in ("mrr", "mtt", "mpp"))
        mrt, mrp, mtp = (float(v[k]) for k in ("mrt", "mrp", "mtp"))
        s = 1e-7
        hdur = float(v["half duration"])
        if hdur < 5.0 * DT:
            hdur = 5.0 * DT
        out.append(CMTSource(x=float(v["longorutm"]), y=float(v["latorutm"]), depth_km=float(v["depth"]),
                             tshift=float(v["time shift"]), hdur=hdur,
                             Mxx=mpp * s, Myy=mtt * s, Mzz=mrr * s, Mxy=-mtp * s, Mxz=mrp * s, Myz=-mrt * s,
                             name=v.get("event name", "")))
    if len(out) == 1:
        out[0].tshift = 0.0
    else:
        m = min(s.tshift for s in out)
        for s in out:
            s.tshift -= m
    return out


@dataclass
class ForceSource:
    """One FORCESOLUTION block (``get_force.f90``): position, STF type and force vector."""
    x: float                    # longorUTM
    y: float                    # latorUTM
    depth: float                # km below the surface, or the z coordinate (m) with USE_SOURCES_RECEIVERS_Z
    tshift: float
    hdur: float                 # half duration (Gaussian, Heaviside) or dominant frequency (Ricker)
    force_stf: int              # 0 Gaussian, 1 Ricker, 2 Heaviside
    factor: float               # factor force source (N)
    e: float                    # component dir vect source E, N, Z_UP
    n: float
    z_up: float


def _fortran_float(v: str) -> float:
    return float(v.strip().lower().replace("d", "e"))


def read_forcesolution(path) -> list[ForceSource]:
    """Parse a FORCESOLUTION file (path or file object; one 11-line block per force)."""
    lines = [l for l in _lines(path) if l.strip()]
    out = []
    for b in range(0, len(lines), 11):
        vals = {}
        for l in lines[b + 1:b + 11]:
            k, v = l.split(":", 1)
            vals[k.strip().lower()] = v.strip()
        out.append(ForceSource(x=_fortran_float(vals["longorutm"]), y=_fortran_float(vals["latorutm"]),
                               depth=_fortran_float(vals["depth"]), tshift=_fortran_float(vals["time shift"]),
                               hdur=_fortran_float(vals["hdurorf0"]), force_stf=int(vals["source time function"]),
                               factor=_fortran_float(vals["factor force source"]),
                               e=_fortran_float(vals["component dir vect source e"]),
                               n=_fortran_float(vals["component dir vect source n"]),
                               z_up=_fortran_float(vals["component dir vect source z_up"])))
    if len(out) == 1:
        out[0].tshift = 0.0
    return out


@dataclass
class Station:
    name: str
    network: str
    y: float    # "latitude" column (UTM y in Cartesian runs)
    x: float    # "longitude" column
    elevation: float
    burial: float


def read_stations(path) -> list[Station]:
    """Parse a STATIONS file (path or file object): ``STA NET lat(y) lon(x) elevation burial``."""
    st = []
    for l in _lines(path):
        p = l.split()
        if len(p) < 6:
            continue
        st.append(Station(p[0], p[1], float(p[2]), float(p[3]), float(p[4]), float(p[5])))
    return st


# ----------------------------------------------------------------------------- location
@dataclass
class Located:
    ispec: int          # 0-based
    xi: float
    eta: float
    gamma: float
    x: float
    y: float
    z: float
    distance: float


class Locator:
    """Point location in a database, following ``locate_point_in_mesh``."""

    NUM_ITER = 5

    def __init__(self, db: Database):
        self.db = db
        self.mid = db.element_midpoints()
        self.xyz = np.stack([db.xstore, db.ystore, db.zstore], 1).astype(np.float64)
        self.anchors = db.anchors
        gll = get_gll()
        self.xigll = gll.xigll
        # element adjacency by shared corner points (for the neighbour retry)
        self._corner_to_elems = None
        self.zmax_surface = float(db.zstore.max())

    def _map(self, ispec, xi, eta, gamma):
        N, dN = hex8_shape(xi, eta, gamma)
        a = self.anchors[ispec]
        x = N @ a
        J = dN @ a  # J[d, c] = d x_c / d xi_d
        Jinv = np.linalg.inv(J.T)  # Jinv[d, c] = d xi_d / d x_c
        return x, Jinv

    def _newton(self, target, ispec, i0, j0, k0):
        xi = self.xigll[i0]; eta = self.xigll[j0]; gamma = self.xigll[k0]
        dmin = 1e30
        for _ in range(self.NUM_ITER):
            x, Jinv = self._map(ispec, xi, eta, gamma)
            d = -(x - target)
            dsq = float(d @ d)
            if dsq < dmin:
                dmin = dsq
                dxi, deta, dgam = Jinv @ d
            else:
                break
            if dxi * dxi + deta * deta + dgam * dgam > 1.0:
                dxi *= 0.33333333333; deta *= 0.33333333333; dgam *= 0.33333333333
            xi = min(max(xi + dxi, -1.01), 1.01)
            eta = min(max(eta + deta, -1.01), 1.01)
            gamma = min(max(gamma + dgam, -1.01), 1.01)
        x, _ = self._map(ispec, xi, eta, gamma)
        return xi, eta, gamma, x

    def _neighbors(self, ispec):
        if self._corner_to_elems is None:
            corners = self.db.ibool[:, [0, 0, 0, 0, 4, 4, 4, 4], [0, 0, 4, 4, 0, 0, 4, 4], [0, 4, 0, 4, 0, 4, 0, 4]]
            from collections import defaultdict
            m = defaultdict(list)
            for e in range(self.db.nspec):
                for g in corners[e]:
                    m[int(g)].append(e)
            self._corner_to_elems = (corners, m)
        corners, m = self._corner_to_elems
        s = set()
        for g in corners[ispec]:
            s.update(m[int(g)])
        s.discard(ispec)
        return sorted(s)

    def locate(self, x, y, z, tie_break: str = "upper") -> Located:
        target = np.array([x, y, z], np.float64)
        d2 = ((self.mid - target) ** 2).sum(1)
        dmin = float(d2.min())
        # Points on element corners/edges/faces are equidistant from several element midpoints;
        # the reference's pick among them depends on its element numbering (kd-tree order and,
        # for NPROC > 1, the partition). tie_break="upper" takes the highest midpoint, then the
        # smallest x, then the smallest y (what the serial internal-mesher run does for a source
        # on a corner of the 20^3 box); "lower" takes the lowest midpoint first.
        tie = np.nonzero(d2 <= dmin * (1.0 + 1e-9) + 1e-12)[0]
        if len(tie) > 1:
            m = self.mid[tie]
            zkey = -m[:, 2] if tie_break == "upper" else m[:, 2]
            order = np.lexsort((m[:, 1], m[:, 0], zkey))
            ispec = int(tie[order[0]])
        else:
            ispec = int(tie[0])
        pts = self.xyz[self.db.ibool[ispec].ravel()]
        dd = ((pts - target) ** 2).sum(1)
        p = int(np.argmin(dd))
        k0, j0, i0 = p // 25, (p // 5) % 5, p % 5
        xi, eta, gamma, xf = self._newton(target, ispec, i0, j0, k0)
        best = (float(((xf - target) ** 2).sum()), ispec, xi, eta, gamma, xf)
        if not (abs(xi) < 1 and abs(eta) < 1 and abs(gamma) < 1):
            # adjacent-element retry: keep the element with the smallest final distance
            # that contains the point (|xi|,|eta|,|gamma| <= 1), as find_best_neighbor does
            for e in self._neighbors(ispec):
                pts = self.xyz[self.db.ibool[e].ravel()]
                dd = ((pts - target) ** 2).sum(1)
                p = int(np.argmin(dd))
                k1, j1, i1 = p // 25, (p // 5) % 5, p % 5
                xi1, eta1, ga1, xf1 = self._newton(target, e, i1, j1, k1)
                dist = float(((xf1 - target) ** 2).sum())
                inside = abs(xi1) <= 1 and abs(eta1) <= 1 and abs(ga1) <= 1
                binside = abs(best[2]) <= 1 and abs(best[3]) <= 1 and abs(best[4]) <= 1
                if (inside and not binside) or (inside == binside and dist < best[0]):
                    best = (dist, e, xi1, eta1, ga1, xf1)
        dist, ispec, xi, eta, gamma, xf = best
        return Located(ispec, xi, eta, gamma, float(xf[0]), float(xf[1]), float(xf[2]), math.sqrt(dist))


def surface_z(db: Database, x: float, y: float) -> float:
    """Elevation of the mesh top surface at (x, y) (flat-top meshes: the max z)."""
    return float(db.zstore.max())


# ----------------------------------------------------------------------------- arrays
def interp_inverse_jacobian(db: Database, ispec: int, hx, hy, hz):
    """Interpolated inverse-Jacobian terms at a point (``compute_arrays_source.f90:76-108``)."""
    # the reference forms hlagrange = hxis(k) * hetas(l) * hgammas(m) in that order
    hl = ((hx[None, None, :] * hy[None, :, None]) * hz[:, None, None]).ravel()
    irr = int(db.irregular_element_number[ispec])
    d = np.zeros(9)
    if irr == 0:
        xr = float(db.xix_regular)
        s = 0.0
        for v in hl:
            s = s + v * xr
        d[0] = d[4] = d[8] = s
    else:
        terms = db.jac9[irr - 1].astype(np.float64)  # (9,125)
        for t in range(9):
            s = 0.0
            for p in range(125):
                s = s + hl[p] * terms[t, p]
            d[t] = s
    return d  # [dxi/dx, dxi/dy, dxi/dz, deta/dx, ..., dgamma/dz]


def cmt_source_array(db: Database, loc: Located, src: CMTSource) -> np.ndarray:
    """(125, 3) float32 source array [k,j,i flattened, comp] (``compute_arrays_source_cmt``)."""
    x = get_gll().xigll
    hx, hpx = lagrange_any(loc.xi, x)
    hy, hpy = lagrange_any(loc.eta, x)
    hz, hpz = lagrange_any(loc.gamma, x)
    d = interp_inverse_jacobian(db, loc.ispec, hx, hy, hz)
    dxis_dx, dxis_dy, dxis_dz, detas_dx, detas_dy, detas_dz, dgs_dx, dgs_dy, dgs_dz = d
    out = np.zeros((5, 5, 5, 3))
    M = src
    for m in range(5):
        for l in range(5):
            for k in range(5):
                hxi = hpx[k] * hy[l] * hz[m]
                heta = hx[k] * hpy[l] * hz[m]
                hga = hx[k] * hy[l] * hpz[m]
                ddx = hxi * dxis_dx + heta * detas_dx + hga * dgs_dx
                ddy = hxi * dxis_dy + heta * detas_dy + hga * dgs_dy
                ddz = hxi * dxis_dz + heta * detas_dz + hga * dgs_dz
                out[m, l, k, 0] = M.Mxx * ddx + M.Mxy * ddy + M.Mxz * ddz
                out[m, l, k, 1] = M.Mxy * ddx + M.Myy * ddy + M.Myz * ddz
                out[m, l, k, 2] = M.Mxz * ddx + M.Myz * ddy + M.Mzz * ddz
    return out.reshape(125, 3).astype(np.float32)


def acoustic_source_array(loc: Located, factor: float) -> np.ndarray:
    """(125, 3) float32 source array of a source in a fluid element: ``factor h_i h_j h_k`` in all three
    components (``compute_arrays_source_forcesolution`` with comp_x = comp_y = comp_z = 1 and the identity
    ``nu_source`` of Cartesian meshes, ``setup_sources_receivers.f90:1150-1212``); only the first
    component is used."""
    x = get_gll().xigll
    hx, _ = lagrange_any(loc.xi, x)
    hy, _ = lagrange_any(loc.eta, x)
    hz, _ = lagrange_any(loc.gamma, x)
    h = ((hx[None, None, :] * hy[None, :, None]) * hz[:, None, None]).ravel()
    v = (factor * h) * 1.0
    return np.repeat(v[:, None], 3, axis=1).astype(np.float32)


def cmt_acoustic_factor(src: "CMTSource") -> float:
    """Scalar moment of a CMT source in a fluid element scaled to the equivalent explosion
    (``setup_sources_receivers.f90:1190-1198``): M0 sqrt(2)/sqrt(3), M0 = sqrt(sum M_ij^2 / 2)."""
    m0 = 1.0 / math.sqrt(2.0) * math.sqrt(src.Mxx ** 2 + src.Myy ** 2 + src.Mzz ** 2
                                          + 2.0 * (src.Myz ** 2 + src.Mxz ** 2 + src.Mxy ** 2))
    return m0 * math.sqrt(2.0) / math.sqrt(3.0)


def force_source_array(loc: Located, fx: float, fy: float, fz: float, factor: float) -> np.ndarray:
    """Point-force source array ``factor * f * h_i h_j h_k`` (float64 -> float32)."""
    x = get_gll().xigll
    hx, _ = lagrange_any(loc.xi, x)
    hy, _ = lagrange_any(loc.eta, x)
    hz, _ = lagrange_any(loc.gamma, x)
    h = ((hx[None, None, :] * hy[None, :, None]) * hz[:, None, None]).ravel()
    f = np.array([fx, fy, fz]) * factor
    return (h[:, None] * f[None, :]).astype(np.float32)


@dataclass
class SourceSet:
    """Everything the loop needs about the sources: arrays + STF per step.

    ``kind`` is the STF of sources in elastic elements (``get_stf_viscoelastic``) and ``kind_acoustic``
    that of sources in fluid elements (``get_stf_acoustic``: a Gaussian, or with the pressure trick its
    second derivative, for CMT sources without Ricker); ``acoustic[s]`` tells which applies."""
    ispec: np.ndarray               # (nsrc,) int32 0-based
    sourcearrays: np.ndarray        # (nsrc, 125, 3) float32
    kind: str                       # "ricker" | "heaviside" | "gaussian"
    hdur: np.ndarray                # (nsrc,) f0 for ricker, hdur otherwise
    tshift: np.ndarray
    t0: float
    DT: float
    located: list = field(default_factory=list)
    acoustic: np.ndarray | None = None   # (nsrc,) bool: source element in a fluid
    kind_acoustic: str = "gaussian"      # "gaussian" | "d2gaussian" | "ricker" | "d2ricker"
    kinds: list | None = None            # per-source STF kinds (point forces), overriding the two above

    @property
    def n(self) -> int:
        return int(self.ispec.shape[0])

    def _kinds(self):
        if self.kinds is not None:
            return list(self.kinds)
        ac = self.acoustic if self.acoustic is not None else np.zeros(self.n, bool)
        return [self.kind_acoustic if a else self.kind for a in ac]

    def stf(self, it: int) -> np.ndarray:
        """float64 STF values at step ``it`` (1-based), ``t = (it-1) DT - t0 - tshift``."""
        t = (it - 1.0) * self.DT - self.t0 - self.tshift
        return np.array([_STF[k](t[s], self.hdur[s]) for s, k in enumerate(self._kinds())], np.float64)

    def stf_table(self, nstep: int) -> np.ndarray:
        """(nstep, nsrc) float64 table of STF values for it = 1..nstep."""
        its = np.arange(1, nstep + 1, dtype=np.float64)
        t = (its[:, None] - 1.0) * self.DT - self.t0 - self.tshift[None, :]
        out = np.empty((nstep, self.n), np.float64)
        for s, k in enumerate(self._kinds()):
            if k not in _STF:
                raise ValueError(k)
            out[:, s] = _STF[k](t[:, s], self.hdur[s])
        return out


def setup_cmt_sources(db: Database, cmts: list[CMTSource], DT: float, ricker: bool = False,
                      tie_break: str = "upper", pressure_trick: bool = False) -> SourceSet:
    """Locate CMT sources and build source arrays, STF kinds and t0.

    Sources in elastic elements get ``compute_arrays_source_cmt``; sources in fluid elements the
    scalar-moment pressure source (``setup_sources_receivers.f90:1188-1212``) with the Gaussian STF (or
    Ricker), or its second derivative when ``pressure_trick`` (USE_TRICK_FOR_BETTER_PRESSURE) is set.
    t0 (``setup_stf_constants``, ``setup_sources_receivers.f90:391-520``): -1.2 min(tshift - 1/f0) for
    Ricker, else -2 min(tshift - hdur), raised to -3 (tshift - hdur) by a Gaussian source in a fluid."""
    loc = Locator(db)
    isp, arrs, located, acoustic = [], [], [], []
    for s in cmts:
        ztop = surface_z(db, s.x, s.y)
        L = loc.locate(s.x, s.y, ztop - s.depth_km * 1000.0, tie_break=tie_break)
        isp.append(L.ispec)
        ac = bool(db.ispec_is_acoustic[L.ispec])
        acoustic.append(ac)
        arrs.append(acoustic_source_array(L, cmt_acoustic_factor(s)) if ac else cmt_source_array(db, L, s))
        located.append(L)
    hdur = np.array([s.hdur for s in cmts], np.float64)
    tshift = np.array([s.tshift for s in cmts], np.float64)
    acoustic = np.array(acoustic, bool)
    if ricker:
        t0 = -1.2 * float(np.min(tshift - 1.0 / hdur))
        kind = "ricker"
        kind_ac = "d2ricker" if pressure_trick else "ricker"
    else:
        t0 = -2.0 * float(np.min(tshift - hdur))
        kind = "heaviside"
        kind_ac = "d2gaussian" if pressure_trick else "gaussian"
        for s in np.nonzero(acoustic)[0]:
            t0 = max(t0, -3.0 * (tshift[s] - hdur[s]))
    return SourceSet(np.array(isp, np.int32), np.stack(arrs), kind, hdur, tshift, t0, DT, located,
                     acoustic=acoustic, kind_acoustic=kind_ac)


def setup_force_sources(db: Database, forces: list[ForceSource], DT: float, ricker: bool = False,
                        z_coordinates: bool = False, pressure_trick: bool = False) -> SourceSet:
    """Point forces (USE_FORCE_POINT_SOURCE): located like CMT sources (``z_coordinates``: the depth
    field is the z coordinate, USE_SOURCES_RECEIVERS_Z); source arrays
    (``setup_sources_receivers.f90:1100-1160``): in elastic elements factor h_ijk times the normalised
    (E, N, Z_UP) vector, in fluid elements factor h_ijk; STF per ``force_stf`` (Ricker for all when
    ``ricker``, as USE_RICKER_TIME_FUNCTION does in get_force.f90:223), with the pressure trick's second
    derivatives in fluids; t0 (``setup_stf_constants``): -min over sources of 1.5 (tshift - hdur)
    (Gaussian, Heaviside) or 1.2 (tshift - 1/f0) (Ricker), raised to -3 (tshift - hdur) by a Gaussian
    force in a fluid."""
    loc = Locator(db)
    isp, arrs, located, acoustic, kinds = [], [], [], [], []
    t0 = 0.0
    for f in forces:
        z = f.depth if z_coordinates else surface_z(db, f.x, f.y) - f.depth * 1000.0
        L = loc.locate(f.x, f.y, z)
        ac = bool(db.ispec_is_acoustic[L.ispec])
        stf = 1 if ricker else f.force_stf
        if ac:
            arr = acoustic_source_array(L, f.factor)
        else:
            norm = math.sqrt(f.e ** 2 + f.n ** 2 + f.z_up ** 2)
            if norm < 1e-30:
                raise ValueError("force point source: component vector has (almost) zero norm")
            arr = force_source_array(L, f.e / norm, f.n / norm, f.z_up / norm, f.factor)
        base = {0: "gaussian", 1: "ricker", 2: "heaviside"}.get(stf)
        if base is None:
            raise ValueError(f"unsupported force source time function {stf}")
        kinds.append(("d2" + base) if (ac and pressure_trick and base != "heaviside") else base)
        t0 = min(t0, 1.2 * (f.tshift - 1.0 / f.hdur) if stf == 1 else 1.5 * (f.tshift - f.hdur))
        isp.append(L.ispec)
        arrs.append(arr)
        located.append(L)
        acoustic.append(ac)
    t0 = -t0
    for f, ac, k in zip(forces, acoustic, kinds):
        if ac and k in ("gaussian", "d2gaussian"):
            t0 = max(t0, -3.0 * (f.tshift - f.hdur))
    hdur = np.array([f.hdur for f in forces], np.float64)
    tshift = np.array([f.tshift for f in forces], np.float64)
    return SourceSet(np.array(isp, np.int32), np.stack(arrs), kinds[0], hdur, tshift, t0, DT, located,
                     acoustic=np.array(acoustic, bool), kinds=kinds)


@dataclass
class ReceiverSet:
    ispec: np.ndarray           # (nrec,) int32 0-based
    hxir: np.ndarray            # (nrec, 5) float64
    hetar: np.ndarray
    hgammar: np.ndarray
    nu: np.ndarray              # (nrec, 3, 3) float64
    names: list
    located: list = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.ispec.shape[0])


def setup_receivers(db: Database, stations: list[Station], z_coordinates: bool = False) -> ReceiverSet:
    """Locate receivers (buried by ``burial`` below the top surface, or at z = ``burial`` with
    ``z_coordinates``, USE_SOURCES_RECEIVERS_Z) and their interpolants."""
    loc = Locator(db)
    x = get_gll().xigll
    isp, hx, hy, hz, names, located = [], [], [], [], [], []
    for s in stations:
        ztop = surface_z(db, s.x, s.y)
        L = loc.locate(s.x, s.y, s.burial if z_coordinates else ztop - s.burial)
        isp.append(L.ispec)
        # hxir_store etc. are CUSTOM_REAL in the reference (specfem3D_par.F90:181)
        hx.append(lagrange_any(L.xi, x)[0].astype(np.float32).astype(np.float64))
        hy.append(lagrange_any(L.eta, x)[0].astype(np.float32).astype(np.float64))
        hz.append(lagrange_any(L.gamma, x)[0].astype(np.float32).astype(np.float64))
        names.append((s.network, s.name))
        located.append(L)
    n = len(stations)
    nu = np.tile(np.eye(3)[None], (n, 1, 1))
    return ReceiverSet(np.array(isp, np.int32), np.array(hx), np.array(hy), np.array(hz), nu, names, located)
