"""GLL constants against closed forms (define_derivation_matrices.f90, gll_library.f90, lagrange_poly.f90)."""
import math

import numpy as np

from specfem3d_amd.gll import get_gll, lagrange_any


def test_gll_nodes_weights_closed_form():
    g = get_gll()
    r = math.sqrt(3.0 / 7.0)
    np.testing.assert_allclose(g.xigll, [-1.0, -r, 0.0, r, 1.0], atol=1e-15)
    np.testing.assert_allclose(g.wgll, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15)


def test_derivative_matrix_exact_on_polynomials():
    g = get_gll()
    x = g.xigll
    for deg in range(5):
        f = x ** deg
        df = deg * x ** (deg - 1) if deg else np.zeros_like(x)
        np.testing.assert_allclose(g.D @ f, df, atol=1e-12)
    np.testing.assert_allclose(g.D.sum(1), 0.0, atol=1e-13)


def test_float32_tables_and_orientation():
    g = get_gll()
    # hprime[i, l] = derivative of Lagrange polynomial l at node i; hprimewgll carries w_i
    assert g.hprime.dtype == np.float32
    np.testing.assert_array_equal(g.hprimewgll, (g.D * g.wgll[:, None]).astype(np.float32))
    np.testing.assert_allclose(g.hprime[0, 0], -5.0, rtol=0)  # -N(N+1)/4 with N = 4
    np.testing.assert_array_equal(g.wgllwgll_xy, np.outer(g.wgll, g.wgll).astype(np.float32))
    # Fortran column-major view: hprime_xx(i2,i1) = l'_{i1}(x_{i2}) at offset i2 + 5 i1
    f = g.fortran_hprime_xx()
    assert f[1 + 5 * 3] == g.hprime[1, 3]


def test_lagrange_any_interpolates_and_differentiates():
    g = get_gll()
    for xi in (-0.83, 0.1234, 0.97):
        h, hp = lagrange_any(xi, g.xigll)
        assert abs(h.sum() - 1.0) < 1e-14
        assert abs(hp.sum()) < 1e-13
        for deg in range(5):
            assert abs((h * g.xigll ** deg).sum() - xi ** deg) < 1e-13
            d = deg * xi ** (deg - 1) if deg else 0.0
            assert abs((hp * g.xigll ** deg).sum() - d) < 1e-12
    h, _ = lagrange_any(g.xigll[2], g.xigll)
    np.testing.assert_allclose(h, [0, 0, 1, 0, 0], atol=1e-15)
