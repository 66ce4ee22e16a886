"""Closed-form pin of the 3-SLS memory-variable update (compute_element_att_memory.f90:28-156,
get_attenuation_model.f90:464-493) in the oracle's stiffness routine.

The update R <- alpha R + f (beta eps_old + gamma eps_new) is the reference's discretisation of
dR/dt = -R / tau_sigma + f eps(t): over one step the exact solution for a strain that is linear in
time is R(t+dt) = exp(-dt/tau) R(t) + f (B eps(t) + G eps(t+dt)) with B = int_0^dt exp(-u/tau) u/dt du
and G = int_0^dt exp(-u/tau) (1 - u/dt) du. Here those integrals come from numerical quadrature,
independent of the Taylor series the code uses, and the oracle is driven with a prescribed uniform
strain history on one element (u_x = eps0 s(t) x, v_x = eps0 s'(t) x). Its memory variables of
every component and SLS must equal the exact solution. Swapping beta and gamma, applying the
factor of the wrong modulus, or mis-indexing a component or an SLS each fails.
"""
import numpy as np
from scipy.integrate import quad

from specfem3d_amd.mesh.box import build_box, homogeneous


def exact_update_weights(tau, dt):
    B = quad(lambda u: np.exp(-u / tau) * u / dt, 0.0, dt, epsabs=0, epsrel=1e-13)[0]
    G = quad(lambda u: np.exp(-u / tau) * (1.0 - u / dt), 0.0, dt, epsabs=0, epsrel=1e-13)[0]
    return np.exp(-dt / tau), B, G


def test_memory_coefficients_against_exact_integrals():
    from specfem3d_amd.attenuation import memory_values
    dt = np.float32(0.004)
    tau = np.array([0.05, 0.3, 2.0], np.float32)
    a, b, g = memory_values(tau, dt)
    for s in range(3):
        A, B, G = exact_update_weights(float(tau[s]), float(dt))
        x = float(dt / tau[s])
        # Taylor truncations of the reference (4th order in dt/tau; beta's last term differs from the
        # exact series by dt^4/(120 tau^3))
        assert abs(a[s] - A) <= 1e-6 + x ** 5
        assert abs(g[s] - G) <= 1e-6 * dt + dt * x ** 3
        assert abs(b[s] - B) <= 1e-6 * dt + dt * x ** 3 / 50
        # gamma weights the new strain more than beta does (a swap fails this when dt/tau is large)
        assert G > B


def test_oracle_memory_variables_follow_the_ode(oracle_mod):
    from specfem3d_amd.attenuation import setup_attenuation
    L = 1000.0
    db = build_box(1, 1, 1, (0, L, 0, L, -L, 0), material=homogeneous(2500.0, 3000.0, 1700.0), stacey_faces=())
    dt = 0.004
    db.finalize(dt, stacey=False)
    qmu = np.full((db.nspec, 125), 40.0, np.float32)
    att = setup_attenuation(db, qmu, qkappa=np.full((db.nspec, 125), 150.0, np.float32), min_resolved_period=0.5,
                            dt=dt)
    o = oracle_mod.OracleSolver(db, None, None, "f64", attenuation=att)
    dtr = float(o.deltat)
    x = db.xstore.astype(np.float64)
    eps0, om = 1e-4, 2 * np.pi * 1.5
    s = lambda t: np.sin(om * t) ** 2
    ds = lambda t: 2 * om * np.sin(om * t) * np.cos(om * t)
    nstep = 200
    ehist = [0.0]
    for n in range(1, nstep + 1):
        t = n * dtr
        o.displ[:] = 0.0
        o.veloc[:] = 0.0
        o.displ[0::3] = eps0 * s(t) * x
        o.veloc[0::3] = eps0 * ds(t) * x
        o.accel[:] = 0.0
        o.forces_elastic()
        ehist.append(eps0 * (s(t) + dtr * ds(t)))   # strain at u + dt v, uniform
    e = np.array(ehist)
    tau = att.tau_sigma.astype(np.float64)
    mu = att.mustore.reshape(-1).astype(np.float64)
    kap = att.kappastore.reshape(-1).astype(np.float64)
    fcm = att.factor_common.reshape(-1, 3).astype(np.float64)
    fck = att.factor_common_kappa.reshape(-1, 3).astype(np.float64)
    comps = {0: 2.0 / 3.0, 1: -1.0 / 3.0, 5: 1.0}   # e_xx, e_yy (deviatoric) and the trace, per unit eps_xx
    for sls in range(3):
        A, B, G = exact_update_weights(float(tau[sls]), dtr)
        r = 0.0
        for n in range(1, nstep + 1):
            r = A * r + B * e[n - 1] + G * e[n]          # exact for piecewise-linear strain, f = 1
        for c, w in comps.items():
            f = (kap * fck[:, sls]) if c == 5 else (mu * fcm[:, sls])
            got = o.R[c].reshape(-1, 3)[:, sls]
            want = f * w * r
            assert np.abs(want).max() > 0
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-5, (sls, c, err)
        for c in (2, 3, 4):  # no shear strain: shear memory only from the float32 hprime row sums (~1e-7)
            assert np.abs(o.R[c]).max() < 1e-5 * np.abs(mu * fcm[:, sls] * r).max()


def test_oracle_stress_correction_is_minus_memory_sum(oracle_mod):
    """The stress seen by the stiffness is sigma - sum_s R (compute_forces_viscoelastic.F90:645-659):
    sigma_xx -= R_xx + R_tr, sigma_yy -= R_yy + R_tr, sigma_zz += R_xx + R_yy - R_tr, sigma_ij -= R_ij.
    For a uniform stress change dS on one cube the nodal forces are closed-form:
    accel_a,i = -sum_j dS_ij int dN_a/dx_j dV = -sum_j dS_ij (L/2)^2 w w [N_a] on the two faces normal
    to x_j (GLL quadrature integrates these exactly)."""
    from specfem3d_amd.attenuation import setup_attenuation
    from specfem3d_amd.gll import get_gll
    L = 1000.0
    db = build_box(1, 1, 1, (0, L, 0, L, -L, 0), material=homogeneous(2500.0, 3000.0, 1700.0), stacey_faces=())
    db.finalize(0.004, stacey=False)
    att = setup_attenuation(db, np.full((1, 125), 40.0, np.float32), qkappa=np.full((1, 125), 150.0, np.float32),
                            min_resolved_period=0.5, dt=0.004)
    o = oracle_mod.OracleSolver(db, None, None, "f64", attenuation=att)
    rng = np.random.default_rng(5)
    rvals = rng.normal(size=(6, 3)) * 1e3          # uniform memory variables, one per component and SLS
    for c in range(6):
        o.R[c][:] = np.tile(rvals[c], 125)
    x = db.xstore.astype(np.float64)
    o.displ[0::3] = 1e-4 * x                        # some strain so the elastic part is non-trivial
    saved_eps = [e.copy() for e in o.eps]
    o.accel[:] = 0.0
    o.forces_elastic()
    a_att = o.accel.copy()
    for c in range(6):
        o.R[c][:] = 0.0
    for e, s in zip(o.eps, saved_eps):
        e[:] = s
    o.accel[:] = 0.0
    o.forces_elastic()
    d = (a_att - o.accel).reshape(-1, 3)
    r = rvals.sum(1)                                # sum over the SLS: xx, yy, xy, xz, yz, trace
    dS = np.array([[-(r[0] + r[5]), -r[2], -r[3]],
                   [-r[2], -(r[1] + r[5]), -r[4]],
                   [-r[3], -r[4], r[0] + r[1] - r[5]]])
    w = get_gll().wgll.astype(np.float64)
    h = (L / 2) ** 2
    want = np.zeros((db.nglob, 3))
    ib = db.ibool.reshape(125)
    for p in range(125):
        i, j, k = p % 5, (p // 5) % 5, p // 25
        I = np.array([h * w[j] * w[k] * ((i == 4) - (i == 0)),
                      h * w[i] * w[k] * ((j == 4) - (j == 0)),
                      h * w[i] * w[j] * ((k == 4) - (k == 0))])
        want[ib[p]] = -dS @ I
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(d - want).max() <= 1e-5 * scale
