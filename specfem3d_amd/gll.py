"""Gauss-Lobatto-Legendre (GLL) constants for NGLL = 5.

Host-side setup constants that cross the boundary into ``prepare_constants_device``:
GLL nodes/weights, the Lagrange derivative matrix ``hprime``, its weighted form
``hprimewgll`` and the 2-D weight products ``wgllwgll_{xy,xz,yz}``.

Reference semantics followed:
  * ``src/shared/define_derivation_matrices.f90:28-103``: everything is built in
    float64 and cast to float32 (``CUSTOM_REAL``) at the end.
  * ``hprime_xx(i2,i1) = l'_{i1}(xi_{i2})`` and
    ``hprimewgll_xx(i2,i1) = l'_{i1}(xi_{i2}) * w_{i2}``
    (``define_derivation_matrices.f90:60-66``), i.e. in 0-based numpy terms
    ``hprime[i, l]`` is the derivative of Lagrange polynomial ``l`` at node ``i``.
  * Lagrange derivative at GLL nodes: ``src/shared/lagrange_poly.f90:208-242``.
  * GLL nodes/weights (Jacobi alpha=beta=0): ``src/shared/gll_library.f90:521-593``.

The arrays here are independent of the element, so they are computed once.
"""
from __future__ import annotations

import numpy as np

NGLLX = NGLLY = NGLLZ = 5
NGLL = 5
NGLL2 = 25
NGLL3 = 125


def _legendre(n: int, x: np.ndarray) -> np.ndarray:
    """Legendre polynomial P_n(x) by the three-term recurrence (float64)."""
    x = np.asarray(x, dtype=np.float64)
    p0 = np.ones_like(x)
    if n == 0:
        return p0
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


def gll_points_weights(n: int = NGLL) -> tuple[np.ndarray, np.ndarray]:
    """GLL nodes (ascending, including +-1) and weights, float64.

    Interior nodes are the roots of P'_{n-1}; weights are 2 / (N (N+1) P_N(x)^2)
    with N = n - 1 (textbook Gauss-Lobatto quadrature).
    """
    N = n - 1
    # roots of P'_N via numpy's Legendre class (exact to ~1e-16 for N=4)
    c = np.zeros(N + 1)
    c[N] = 1.0
    dP = np.polynomial.legendre.legder(c)
    interior = np.sort(np.polynomial.legendre.legroots(dP).real)
    x = np.concatenate(([-1.0], interior, [1.0]))
    # symmetrise exactly (the reference's zwgljd yields symmetric nodes)
    x = 0.5 * (x - x[::-1])
    if n % 2 == 1:
        x[n // 2] = 0.0
    w = 2.0 / (N * (N + 1) * _legendre(N, x) ** 2)
    return x, w


def lagrange_deriv_matrix(x: np.ndarray) -> np.ndarray:
    """D[i, l] = l'_l(x_i), the derivative of Lagrange polynomial l at node i.

    Closed form at GLL nodes (``lagrange_poly.f90:208-242``):
    D[0,0] = -N(N+1)/4, D[N,N] = N(N+1)/4, other diagonal 0,
    D[i,l] = P_N(x_i) / (P_N(x_l) (x_i - x_l)) off the diagonal.
    """
    n = len(x)
    N = n - 1
    P = _legendre(N, x)
    D = np.zeros((n, n))
    for i in range(n):
        for l in range(n):
            if i == l:
                if i == 0:
                    D[i, l] = -N * (N + 1) * 0.25
                elif i == N:
                    D[i, l] = N * (N + 1) * 0.25
            else:
                D[i, l] = P[i] / (P[l] * (x[i] - x[l]))
    return D


def lagrange_any(xi: float, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Lagrange interpolants h_l(xi) and derivatives h'_l(xi), float64.

    Restates ``src/shared/lagrange_poly.f90`` ``lagrange_any`` (product form) for an
    arbitrary point; used for source arrays and receiver interpolation.
    """
    n = len(x)
    h = np.empty(n)
    hp = np.empty(n)
    for d in range(n):
        prod1 = 1.0
        prod2 = 1.0
        for i in range(n):
            if i != d:
                prod1 *= xi - x[i]
                prod2 *= x[d] - x[i]
        inv = 1.0 / prod2
        h[d] = prod1 * inv
        s = 0.0
        for i in range(n):
            if i != d:
                p3 = 1.0
                for j in range(n):
                    if j != d and j != i:
                        p3 *= xi - x[j]
                s += p3
        hp[d] = s * inv
    return h, hp


class GLL:
    """All NGLL=5 constants, float64 originals and float32 device copies."""

    def __init__(self):
        x, w = gll_points_weights(NGLL)
        self.xigll = x
        self.wgll = w
        D = lagrange_deriv_matrix(x)
        self.D = D
        # float32 arrays exactly as the reference casts them
        self.hprime = D.astype(np.float32)                          # [i, l] = l'_l(x_i)
        self.hprimewgll = (D * w[:, None]).astype(np.float32)       # [i, l] = l'_l(x_i) w_i
        ww = np.outer(w, w)
        self.wgllwgll_xy = ww.astype(np.float32)                    # [i, j] = w_i w_j
        self.wgllwgll_xz = ww.astype(np.float32)                    # [i, k]
        self.wgllwgll_yz = ww.astype(np.float32)                    # [j, k]

    # Fortran column-major views the C-ABI receives (hprime_xx(i2,i1) = l'_{i1}(x_{i2}))
    def fortran_hprime_xx(self) -> np.ndarray:
        """Column-major flat copy of hprime_xx(i2,i1) as the Fortran host passes it."""
        return np.asfortranarray(self.hprime).ravel(order="F").astype(np.float32)

    def fortran_hprimewgll_xx(self) -> np.ndarray:
        return np.asfortranarray(self.hprimewgll).ravel(order="F").astype(np.float32)

    def fortran_wgllwgll(self) -> np.ndarray:
        return np.asfortranarray(self.wgllwgll_xy).ravel(order="F").astype(np.float32)


_GLL = None


def get_gll() -> GLL:
    global _GLL
    if _GLL is None:
        _GLL = GLL()
    return _GLL
