"""Standard-linear-solid (N_SLS = 3) attenuation set-up on the host.

Produces the inputs of the viscoelastic stiffness kernel -- ``factor_common``,
``factor_common_kappa`` (N_SLS per GLL point), the physically dispersed ``mustore`` /
``kappastore`` and the Runge-Kutta coefficients ``alphaval/betaval/gammaval`` -- following
``src/shared/get_attenuation_model.f90``:

* band: ``MAX_PERIOD = MIN_PERIOD * 10**1.75`` for 3 SLS (``get_attenuation_periods`` :702-748);
* relaxation times tau_sigma log-spaced over the band (``get_attenuation_tau_sigma`` :754-785),
  centre frequency ``f_c`` the geometric mean (:791-810);
* tau_epsilon per Q fitted by a Nelder-Mead simplex search of the summed relative misfit of
  ``tan(delta) = B/A`` of the Maxwell bodies against ``1/Q`` at 100 log-spaced frequencies
  (``attenuation_invert_by_simplex``/``attenuation_eval``/``attenuation_maxwell`` :976-1215),
  started at ``tau_eps = tau_s (1 + 2/Q)`` and cached per Q rounded to 0.1 (:879-970). The search
  is the reference's own ``fminsearch`` (:1222-1607) step for step -- coefficients, the 5 % initial
  simplex, the absolute tolerances 1e-4 on tau and misfit, 200 n iterations, the stable bubble sort
  of the vertices -- because with an absolute tolerance of 1e-4 on tau values of 3e-3 the point it
  stops at depends on the path, not only on the minimum;
* ``beta = tau_eps/tau_s - 1``, ``one_minus_sum_beta = sum tau_eps/tau_s`` (:599-628);
* modulus scale factor to the reference frequency (``get_attenuation_scale_factor`` :634-694);
* ``factor_common = 2 beta / tau_s / one_minus_sum_beta`` (mu) and ``beta / tau_s / ...`` (kappa)
  (:405-415); mustore/kappastore multiplied by the scale factors (``prepare_attenuation.f90``);
* RK coefficients = 4th-order Taylor series of exp(-dt/tau_s) (``get_attenuation_memory_values``
  :464-493), evaluated in float32.

The reference's viscoelastic benchmark (EXAMPLES/benchmarks/attenuation/viscoelastic) lists the
tau_sigma, tau_epsilon and scale factors its analytical solution was computed with (Q_kappa = 10,
Q_mu = 20, f0 = 18 Hz on its mesh); these routines reproduce them (tests/test_attenuation_reference.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

N_SLS = 3
ATTENUATION_COMP_MAXIMUM = 9000.0
THETA = {2: 0.75, 3: 1.75, 4: 2.25, 5: 2.85}


def attenuation_periods(min_resolved_period: float) -> tuple[float, float]:
    tmin = float(min_resolved_period)
    return tmin, tmin * 10.0 ** THETA[N_SLS]


def tau_sigma(min_period: float, max_period: float) -> np.ndarray:
    PI = 3.14159265358979
    e1 = math.log10(1.0 / max_period)
    e2 = math.log10(1.0 / min_period)
    d = (e2 - e1) / (N_SLS - 1.0)
    return np.array([1.0 / (PI * 2.0 * 10 ** (e1 + i * d)) for i in range(N_SLS)])


def source_freq(min_period: float, max_period: float) -> float:
    return 10.0 ** (0.5 * (math.log10(1.0 / max_period) + math.log10(1.0 / min_period)))


def _maxwell(f, tau_s, tau_eps):
    """A, B of the N_SLS Maxwell bodies at log10-frequencies f (``attenuation_maxwell``), each
    accumulated over the bodies in order."""
    PI = 3.14159265358979
    w = 2.0 * PI * 10.0 ** f
    A = np.zeros_like(w)
    B = np.zeros_like(w)
    for j in range(len(tau_s)):
        den = 1.0 + w ** 2 * tau_s[j] ** 2
        A += (1.0 + (w ** 2 * tau_eps[j] * tau_s[j])) / den
        B += w * (tau_eps[j] - tau_s[j]) / den
    return A, B


# REAL (single precision) literals of the reference's fminsearch, as the double values they become
_F32 = lambda v: float(np.float32(v))


def fminsearch(funk, x0: np.ndarray):
    """Nelder-Mead search as ``fminsearch`` (get_attenuation_model.f90:1222-1473) with the tolerances
    ``attenuation_invert_by_simplex`` passes (absolute 1e-4 on the simplex size and on the spread of
    the function values, 200 n iterations and evaluations). Returns (x, f(x), iterations, evaluations)."""
    n = len(x0)
    rho, chi, psi, sigma = 1.0, 2.0, 0.5, 0.5
    maxiter = maxfun = 200 * n
    tolx = tolf = _F32(1.0e-4)
    usual_delta = _F32(0.05)
    zero_term_delta = _F32(0.00025)
    xin = np.asarray(x0, np.float64).copy()
    v = [xin.copy()]
    fv = [funk(xin)]
    for j in range(n):
        y = xin.copy()
        y[j] = (1.0 + usual_delta) * y[j] if y[j] != 0.0 else zero_term_delta
        v.append(y)
        fv.append(funk(y))

    def sort():
        # qsort_local: bubble sort on the values, stable (swaps only on a strict decrease)
        idx = list(range(n + 1))
        f = list(fv)
        for j in range(n + 1):
            for k in range(n - j):
                if f[k + 1] < f[k]:
                    f[k], f[k + 1] = f[k + 1], f[k]
                    idx[k], idx[k + 1] = idx[k + 1], idx[k]
        return [v[i] for i in idx], f

    v, fv = sort()
    itercount, func_evals = 1, n + 1
    while func_evals < maxfun and itercount < maxiter:
        size = max(abs(v[j][i] - v[0][i]) for i in range(n) for j in range(1, n + 1))
        spread = max(abs(fv[0] - fv[i]) for i in range(1, n + 1))
        if size <= tolx and spread <= tolf:
            break
        xbar = np.zeros(n)
        for i in range(n):
            acc = 0.0
            for j in range(n):
                acc = acc + v[j][i]
            xbar[i] = acc / (n * 1.0)
        xr = (1 + rho) * xbar - rho * v[n]
        fxr = funk(xr)
        func_evals += 1
        if fxr < fv[0]:
            xe = (1 + rho * chi) * xbar - rho * chi * v[n]
            fxe = funk(xe)
            func_evals += 1
            if fxe < fxr:
                v[n], fv[n] = xe, fxe
            else:
                v[n], fv[n] = xr, fxr
        elif fxr < fv[n - 1]:
            v[n], fv[n] = xr, fxr
        else:
            shrink = False
            if fxr < fv[n]:
                xc = (1 + psi * rho) * xbar - psi * rho * v[n]
                fxc = funk(xc)
                func_evals += 1
                if fxc <= fxr:
                    v[n], fv[n] = xc, fxc
                else:
                    shrink = True
            else:
                xcc = (1 - psi) * xbar + psi * v[n]
                fxcc = funk(xcc)
                func_evals += 1
                if fxcc < fv[n]:
                    v[n], fv[n] = xcc, fxcc
                else:
                    shrink = True
            if shrink:
                for j in range(1, n + 1):
                    v[j] = v[0] + sigma * (v[j] - v[0])
                    fv[j] = funk(v[j])
                func_evals += n
        v, fv = sort()
        itercount += 1
    if func_evals > maxfun or itercount > maxiter:
        raise RuntimeError("attenuation_simplex: search for strain relaxation times did not converge")
    return v[0], fv[0], itercount, func_evals


_TAU_EPS_CACHE: dict = {}


def tau_eps_for_q(Q: float, tau_s: np.ndarray, min_period: float, max_period: float) -> np.ndarray:
    """Strain relaxation times fitting a constant Q over the band (``get_attenuation_tau_eps`` with
    ``attenuation_invert_by_simplex``), cached per int(10 Q) like ``model_attenuation_storage``."""
    key = (int(Q * 10), tuple(np.round(tau_s, 12)), min_period, max_period)
    if key in _TAU_EPS_CACHE:
        return _TAU_EPS_CACHE[key]
    nf = 100
    e1 = math.log10(1.0 / max_period)
    e2 = math.log10(1.0 / min_period)
    f = np.array([e1 + ((i * 1.0) * (e2 - e1) / ((nf - 1) * 1.0)) for i in range(nf)])
    iQ = 1.0 / Q
    iQ2 = iQ ** 2

    def misfit(x):
        A, B = _maxwell(f, tau_s, x)
        xi = np.sqrt(((B / A - iQ) ** 2) / iQ2)
        acc = 0.0
        for val in xi.tolist():     # summed in order, as attenuation_eval
            acc = acc + val
        return acc

    x0 = np.array([t + (t * 2.0 / Q) for t in tau_s], np.float64)
    x, _, _, _ = fminsearch(misfit, x0)
    _TAU_EPS_CACHE[key] = np.asarray(x, np.float64)
    return _TAU_EPS_CACHE[key]


def scale_factor(f_c: float, tau_e: np.ndarray, tau_s: np.ndarray, Q: float, f0_ref: float) -> float:
    PI = math.pi
    s0 = 1.0 + 2.0 * math.log(f_c / f0_ref) / (PI * Q)
    x1 = 1.0
    x2 = 1.0
    for i in range(N_SLS):
        a = tau_e[i] / tau_s[i] - 1.0
        x1 += a / N_SLS
        x2 += a / (1.0 + 1.0 / (2.0 * PI * f_c * tau_s[i]) ** 2) / N_SLS
    s = (x1 / x2) * s0
    if s < 0.5 or s > 1.5:
        raise ValueError("unreliable correction factor in attenuation model")
    return s


def memory_values(tau_s32: np.ndarray, deltat: np.float32):
    """alphaval, betaval, gammaval in float32 (``get_attenuation_memory_values``)."""
    f = np.float32
    dt = f(deltat)
    tauinv = (-(f(1) / tau_s32.astype(f))).astype(f)
    a = (f(1.0) + dt * tauinv + dt ** 2 * tauinv ** 2 / f(2) + dt ** 3 * tauinv ** 3 / f(6)
         + dt ** 4 * tauinv ** 4 / f(24)).astype(f)
    b = (dt / f(2) + dt ** 2 * tauinv / f(3) + dt ** 3 * tauinv ** 2 / f(8) + dt ** 4 * tauinv ** 3 / f(24)).astype(f)
    g = (dt / f(2) + dt ** 2 * tauinv / f(6) + dt ** 3 * tauinv ** 2 / f(24)).astype(f)
    return a, b, g


@dataclass
class Attenuation:
    factor_common: np.ndarray        # (nspec, 125, 3) float32
    factor_common_kappa: np.ndarray
    kappastore: np.ndarray           # (nspec, 125) float32, scaled
    mustore: np.ndarray
    alphaval: np.ndarray
    betaval: np.ndarray
    gammaval: np.ndarray
    tau_sigma: np.ndarray            # float32
    f_c: float
    min_period: float
    max_period: float
    scale_factor: np.ndarray | None = None        # (nspec, 125) float32 modulus scale of mu (1 outside)
    scale_factor_kappa: np.ndarray | None = None  # (nspec, 125) float32 modulus scale of kappa


def setup_attenuation(db, qmu, qkappa=None, min_resolved_period: float | None = None, f0_ref: float = 18.0,
                      dt: float | None = None) -> Attenuation:
    """Build SLS factors for a database from per-GLL-point Q_mu (and Q_kappa) arrays (nspec,125)."""
    if min_resolved_period is None:
        min_resolved_period = estimate_min_resolved_period(db)
    tmin, tmax = attenuation_periods(0.99 * np.float32(min_resolved_period))
    ts = tau_sigma(tmin, tmax)
    fc = source_freq(tmin, tmax)
    ts32 = ts.astype(np.float32)
    tauinv = (np.float32(1) / ts32).astype(np.float32)
    nspec = db.nspec
    fcm = np.ones((nspec, 125, N_SLS), np.float32)
    fck = np.ones((nspec, 125, N_SLS), np.float32)
    sc = np.ones((nspec, 125), np.float32)
    sck = np.ones((nspec, 125), np.float32)
    qmu = np.broadcast_to(np.asarray(qmu, np.float64), (nspec, 125))
    qk = (np.full((nspec, 125), ATTENUATION_COMP_MAXIMUM) if qkappa is None
          else np.broadcast_to(np.asarray(qkappa, np.float64), (nspec, 125)))
    el = db.ispec_is_elastic
    # Q clamping as get_attenuation_model.f90:336-372; points with Q_mu <= 1e-5 keep factor 1
    Qm = qmu[el]
    Qk = qk[el]
    active = Qm > 1e-5
    Qm = np.clip(Qm, 1.0, ATTENUATION_COMP_MAXIMUM)
    Qk = np.where(Qk <= 1e-5, ATTENUATION_COMP_MAXIMUM, Qk)
    Qk = np.clip(Qk, 1.0, ATTENUATION_COMP_MAXIMUM)
    # one fit per distinct (Q_mu, Q_kappa) pair, broadcast to the points
    if Qm.size and (Qm == Qm.flat[0]).all() and (Qk == Qk.flat[0]).all():   # uniform Q: one fit
        pairs = np.array([[Qm.flat[0], Qk.flat[0]]])
        inv = np.zeros(Qm.shape, np.int64)
    else:
        key = Qm.ravel() + 1j * Qk.ravel()                          # 1-D sort, (Q_mu, Q_kappa) order
        uk, inv = np.unique(key, return_inverse=True)
        pairs = np.stack([uk.real, uk.imag], 1)
        inv = inv.reshape(Qm.shape)
    tab_a = np.empty((len(pairs), N_SLS), np.float32)
    tab_b = np.empty((len(pairs), N_SLS), np.float32)
    tab_s = np.empty(len(pairs), np.float32)
    tab_sk = np.empty(len(pairs), np.float32)
    for u, (qm_u, qk_u) in enumerate(pairs):
        te_k = tau_eps_for_q(float(qk_u), ts, tmin, tmax)
        beta_k = te_k / ts
        omsb_k = float(beta_k.sum())
        beta_k = beta_k - 1.0
        te = tau_eps_for_q(float(qm_u), ts, tmin, tmax)
        beta = te / ts
        omsb = float(beta.sum())
        beta = beta - 1.0
        tab_a[u] = ((2.0 * beta * tauinv.astype(np.float64)) / omsb).astype(np.float32)
        tab_b[u] = ((beta_k * tauinv.astype(np.float64)) / omsb_k).astype(np.float32)
        tab_s[u] = np.float32(scale_factor(fc, te, ts, float(qm_u), f0_ref))
        tab_sk[u] = np.float32(scale_factor(fc, te_k, ts, float(qk_u), f0_ref))
    sub_a = np.where(active[..., None], tab_a[inv], np.float32(1))
    sub_b = np.where(active[..., None], tab_b[inv], np.float32(1))
    fcm[el] = sub_a
    fck[el] = sub_b
    sc[el] = np.where(active, tab_s[inv], np.float32(1))
    sck[el] = np.where(active, tab_sk[inv], np.float32(1))
    mu = db.mustore.copy()
    kap = db.kappastore.copy()
    mu[el] = (mu[el] * sc[el]).astype(np.float32)
    kap[el] = (kap[el] * sck[el]).astype(np.float32)
    deltat = np.float32(db.dt if dt is None else dt)
    al, be, ga = memory_values(ts32, deltat)
    return Attenuation(fcm, fck, kap, mu, al, be, ga, ts32, fc, tmin, tmax, sc, sck)


def estimate_min_resolved_period(db) -> float:
    """Shortest period resolved with ~5 points per wavelength at the slowest S (or P) speed."""
    h = float(db.meta.get("hx", 0.0)) or float(np.cbrt(8.0 * float(db.jacobian_regular))) if db.nspec_irregular == 0 else None
    if not h:
        a = db.anchors
        h = float(np.max(np.linalg.norm(a[:, 1] - a[:, 0], axis=-1)))
    rho = db.rhostore.astype(np.float64)
    vs = np.where(db.mustore > 0, np.sqrt(np.maximum(db.mustore, 0) / rho), np.inf)
    vmin = float(np.min(vs[db.ispec_is_elastic])) if db.nspec_elastic else 1.0
    return (h / 4.0) * 5.0 / vmin
