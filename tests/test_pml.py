"""C-PML set-up and CPU oracle (no GPU).

* the reference example CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides: the
  C-PML widths, interface origins and element count that its own generate_databases run printed;
* the recursive-convolution coefficients against the exact convolution of a constant (steady state
  1/b), the exponential, and the Taylor branch against the closed form at small b;
* the acceleration and strain coefficients against the partial-fraction expansions they are
  (s^2 s_x s_y s_z and s_i s_j / s_k with s_i = K_i + d_i / (alpha_i + s)), region by region;
* the C-PML mass matrix: without damping it is the plain mass, and it grows with d;
* the oracle: energy leaves a box through its C-PML layers and stays in the same box without them;
  a layer in place of the free surface (PML_INSTEAD_OF_FREE_SURFACE) takes out the energy the free
  surface keeps, and its damping profile mirrors the bottom layer's.
"""
import numpy as np
import pytest

from conftest import GOLDEN, small_pml_box

from specfem3d_amd import pml as P

f32 = np.float32


def test_cpml5_setup_matches_reference_log():
    from specfem3d_amd.examples import cpml5_example
    db, src, rec, ref, dt, z = cpml5_example()
    p = db.pml
    assert p.n == int(z["log_nspec_cpml"]) == 9036
    np.testing.assert_allclose(p.width, z["log_cpml_width"], rtol=1e-7)
    o = z["log_cpml_origins"]  # right/left x, front/back y, bottom z
    got = [p.origins["x_right"], p.origins["x_left"], p.origins["y_front"], p.origins["y_back"], p.origins["z_bottom"]]
    np.testing.assert_allclose(got, o, rtol=1e-7)
    # damping only inside the layers, zero at the interface, largest at the outer edge
    assert np.all(p.d >= 0) and np.all(p.alpha >= 0) and np.all(p.K == 1)
    assert p.d.max() > 1.0
    assert src.ispec[0] not in set(p.cpml_to_spec.tolist())


@pytest.mark.parametrize("b", [1e-3, 0.05, 0.3, 2.0])
def test_convolution_coefficients_steady_state(b):
    """psi(n+1) = c0 psi(n) + c1 u(n+1) + c2 u(n) with u = 1 converges to int exp(-b t) dt = 1/b, and
    c0 = exp(-b dt) (prepare_timerun.F90:926-1030)."""
    dt = f32(0.01)
    c0, c1, c2 = (float(v) for v in P.convolution_coef(np.array([b], np.float32), dt, f32(1e-6)))
    assert abs(c0 - np.exp(-b * float(dt))) < 2e-7
    psi = (c1 + c2) / (1.0 - c0)
    assert abs(psi * b - 1.0) < 2e-3 if b < 0.01 else abs(psi * b - 1.0) < 2e-4


def test_convolution_taylor_branch():
    """Below the separation distance the coefficients come from third-order Taylor expansions of the
    closed forms (1 - e^{-b dt/2}) / b and e^{-b dt/2} (1 - e^{-b dt/2}) / b."""
    dt = f32(0.11)
    b = np.array([0.0, 1e-4, 1e-3], np.float32)
    _, c1, c2 = P.convolution_coef(b, dt, f32(1.0))
    bb = b.astype(np.float64)
    h = np.exp(-0.5 * bb * float(dt))
    e1 = np.where(bb > 0, (1 - h) / np.where(bb > 0, bb, 1), 0.5 * float(dt))
    np.testing.assert_allclose(c1, e1, rtol=1e-6)
    np.testing.assert_allclose(c2, e1 * h, rtol=1e-6)


def _s(k, d, a, s):
    return k + d / (a + s)


REGIONS = [P.CPML_X_ONLY, P.CPML_Y_ONLY, P.CPML_Z_ONLY, P.CPML_XY_ONLY, P.CPML_XZ_ONLY, P.CPML_YZ_ONLY, P.CPML_XYZ]


def _params(region, rng):
    has = P._HAS[region]
    K = np.where(has, rng.uniform(1.0, 1.5, 3), 1.0).astype(np.float32)
    d = np.where(has, rng.uniform(0.2, 2.0, 3), 0.0).astype(np.float32)
    a = np.where(has, np.array([0.11, 0.23, 0.37]) * rng.uniform(0.8, 1.2, 3), 0.0).astype(np.float32)
    return K, d, a


@pytest.mark.parametrize("region", REGIONS)
def test_acceleration_coefficients_partial_fractions(region):
    """s^2 s_x s_y s_z = A0 s^2 + A1 s + A2 + A3/(s+alpha_x) + A4/(s+alpha_y) + A5/(s+alpha_z), A0 = Kx Ky Kz
    (l_parameter_computation, pml_compute_accel_contribution.f90:445-597)."""
    rng = np.random.default_rng(region)
    K, d, a = _params(region, rng)
    A = [float(np.ravel(v)[0]) for v in P._l_parameters(region, *[np.array([v]) for v in (K[0], d[0], a[0], K[1], d[1], a[1],
                                                                             K[2], d[2], a[2])], f32(1e-4))]
    A0 = float(K[0]) * float(K[1]) * float(K[2])
    Kd, dd, ad = K.astype(float), d.astype(float), a.astype(float)
    for s in (0.3 + 0.7j, 1.7 - 0.2j, 5.0 + 3.0j):
        lhs = s * s * _s(Kd[0], dd[0], ad[0], s) * _s(Kd[1], dd[1], ad[1], s) * _s(Kd[2], dd[2], ad[2], s)
        rhs = A0 * s * s + A[0] * s + A[1] + A[2] / (s + ad[0]) + A[3] / (s + ad[1]) + A[4] / (s + ad[2])
        assert abs(lhs - rhs) <= 2e-5 * abs(lhs), (s, lhs, rhs)


@pytest.mark.parametrize("region", REGIONS)
@pytest.mark.parametrize("index", [123, 132, 231])
def test_strain_coefficients_partial_fractions(region, index):
    """For the permuted arguments (x, y, z) of lijk_parameter_computation (pml_compute_memory_variables.f90:
    912-1070): s_x s_y / s_z = A0 + A1/(s+alpha_x) + A2/(s+alpha_y) + A3/(s+beta_z), beta = alpha + d/K;
    and l{x,y,z}: s_x = A0 + A1/(s+alpha_x)."""
    rng = np.random.default_rng(100 * region + index % 7)
    K, d, a = _params(region, rng)
    perm = {123: (0, 1, 2), 132: (0, 2, 1), 231: (2, 1, 0)}[index]
    args = []
    for ax in perm:
        args += [np.array([K[ax]]), np.array([d[ax]]), np.array([a[ax]])]
    A = [float(np.ravel(v)[0]) for v in P._lijk(*args, region, index, f32(1e-4))]
    Kp, dp, ap = (np.array([v[i] for i in perm], float) for v in (K, d, a))
    bz = ap[2] + dp[2] / Kp[2]
    for s in (0.4 + 0.5j, 2.0 - 1.0j):
        lhs = _s(Kp[0], dp[0], ap[0], s) * _s(Kp[1], dp[1], ap[1], s) / _s(Kp[2], dp[2], ap[2], s)
        rhs = A[0] + A[1] / (s + ap[0]) + A[2] / (s + ap[1]) + A[3] / (s + bz)
        assert abs(lhs - rhs) <= 2e-5 * abs(lhs), (s, lhs, rhs)
    for axis in range(3):
        A0, A1 = (float(np.ravel(v)[0]) for v in P._l_axis(np.array([K[axis]]), np.array([d[axis]]), np.array([a[axis]]),
                                               region, axis))
        for s in (0.4 + 0.5j, 2.0 - 1.0j):
            assert abs(_s(float(K[axis]), float(d[axis]), float(a[axis]), s) - (A0 + A1 / (s + float(a[axis])))) < 1e-5


def test_pml_with_fluid_elements_is_refused():
    """The reference's GPU path stops on C-PML conditions in a mesh with fluid elements
    (compute_forces_acoustic_calling_routine.f90:97); the database does the same."""
    from conftest import small_coupled_box
    db, _, _ = small_coupled_box(n=4)
    cs, rg = P.box_regions(db, 1, ("xmin", "xmax", "ymin", "ymax", "bottom"))
    solid = db.ispec_is_elastic[cs]
    db.pml = P.set_damping(db, cs[solid], rg[solid], f0=4.0)
    with pytest.raises(ValueError, match="fluid elements"):
        db.finalize(db.dt, stacey=False)


def test_pml_mass_matrix():
    db, _, _, _ = small_pml_box(n=4, layers=1)
    p = db.pml
    m_pml = P.pml_elastic_mass(db, p, db.dt)
    d0 = p.d.copy()
    p.d[:] = 0
    m_plain = P.pml_elastic_mass(db, p, db.dt)
    p.d[:] = d0
    from specfem3d_amd.mesh.database import _ordered_scatter_add
    from specfem3d_amd.gll import get_gll
    w = get_gll().wgll
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()
    jac = np.full((db.nspec, 125), db.jacobian_regular, np.float32)
    base = (jac.astype(np.float64) * w3 * db.rhostore.astype(np.float64)).astype(np.float32)
    order = np.concatenate([np.nonzero(~p.is_cpml(db.nspec))[0], p.cpml_to_spec])
    plain = _ordered_scatter_add(db.nglob, db.ibool.reshape(db.nspec, 125)[order], base[order])
    assert np.array_equal(m_plain, plain)
    inside = np.zeros(db.nglob, bool)
    inside[db.ibool.reshape(db.nspec, 125)[p.cpml_to_spec]] = True
    assert np.all(m_pml[~inside] == plain[~inside])
    assert np.all(m_pml[inside] >= plain[inside]) and np.any(m_pml[inside] > plain[inside])


def test_oracle_energy_leaves_through_the_layers(oracle_mod):
    """Kinetic energy after the waves have crossed the box: with C-PML layers it falls well below the
    energy of the same box with traction-free outer faces (no absorbing condition), where it stays."""
    out = {}
    for layers in (0, 2):
        db, src, rec, _ = small_pml_box(n=6, layers=2)
        if layers == 0:
            db.pml = None
            db.finalize(db.dt, stacey=False)
        o = oracle_mod.OracleSolver(db, src, rec)
        mass = 1.0 / db.rmassx.astype(np.float64)
        e = []
        for it in range(900):
            o.step()
            if (it + 1) % 300 == 0:
                v = o.veloc.reshape(-1, 3).astype(np.float64)
                e.append(0.5 * float((mass * (v * v).sum(1)).sum()))
        assert np.isfinite(o.displ).all()
        out[layers] = e
    print("kinetic energy no layers", out[0], "with layers", out[2])
    assert out[2][-1] < 0.2 * out[0][-1]
    assert out[2][-1] < out[2][0]


def test_top_layer_mirrors_bottom_layer():
    """PML_INSTEAD_OF_FREE_SURFACE: the top layer's z profile is measured from the upper interface
    (pml_set_local_dampingcoeff.f90, the z_origin_top branch). On a box with equal top and bottom
    layers, d_z and alpha_z at a point of the top layer equal those at its mirror image in the
    bottom layer, and both vanish at the interfaces and peak on the outer faces."""
    db, _, _, _ = small_pml_box(n=6, layers=2, top=True)
    p = db.pml
    assert p.instead_of_free_surface and p.origins["z_top"] is not None
    ib = db.ibool.reshape(db.nspec, 125)[p.cpml_to_spec]
    z = db.zstore[ib].astype(np.float64)
    zmin, zmax = float(db.zstore.min()), float(db.zstore.max())
    damped = np.array([P._HAS[int(r)][2] for r in p.regions], bool)
    zz, dz, az = z[damped].ravel(), p.d[2][damped].ravel(), p.alpha[2][damped].ravel()
    top = zz > 0.5 * (zmin + zmax)
    assert top.any() and (~top).any()
    # distance into the layer from the nearer outer face, rounded to the mesh's float32 resolution;
    # the points at one distance carry the same values on both sides (alpha also after the separation
    # of nearly equal parameters, which moves some of them: compared as sorted sets)
    key = np.round(np.where(top, zmax - zz, zz - zmin), 3)
    for k in np.unique(key):
        for v in (dz, az):
            a, b = np.sort(v[top & (key == k)]), np.sort(v[~top & (key == k)])
            assert a.size == b.size > 0
            np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-6)
    assert dz[top].max() == pytest.approx(dz.max(), rel=2e-5) and dz[top & (key == key.max())].max() == 0
    assert np.all(dz[top & (key == 0)] == dz[top].max())


def test_oracle_top_layer_takes_out_surface_energy(oracle_mod):
    """Kinetic energy 900 steps after the source: with a C-PML layer in place of the free surface it
    falls far below the five-sided box's, where the free surface keeps waves in the box."""
    out = {}
    for top in (False, True):
        db, src, rec, _ = small_pml_box(n=6, layers=2, top=top)
        o = oracle_mod.OracleSolver(db, src, rec)
        mass = 1.0 / db.rmassx.astype(np.float64)
        for _ in range(900):
            o.step()
        v = o.veloc.reshape(-1, 3).astype(np.float64)
        out[top] = 0.5 * float((mass * (v * v).sum(1)).sum())
        assert np.isfinite(o.displ).all()
    print("kinetic energy five sides", out[False], "six sides", out[True])
    assert out[True] < 0.05 * out[False]
