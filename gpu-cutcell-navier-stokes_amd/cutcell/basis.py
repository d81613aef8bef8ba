"""Open-uniform (clamped) B-spline bases on [0, 1] and their tables.

This is the host-side input layer of the cut-cell PCG path: every system the
solver sees is a Galerkin matrix over tensor products of these 1D bases.

Reference behaviour followed (file:line in the reference repository):
- knot vector, Greville abscissae and supports: ``code/cut_spline_sbp.py:52-76``;
- monomial spline coefficients by Greville collocation:
  ``code/cut_spline_sbp.py:148-153``;
- vectorised value / first-derivative tables through
  ``scipy.interpolate.BSpline.design_matrix`` with the degree-lowering
  derivative recurrence, and local element support ``[e, e+p+1)``:
  ``code/cut_assembly.py:31-78``.

The Greville abscissae are computed with the same floating-point operations
(an ``np.sum`` of ``p`` consecutive knots, then one division) because the
aggregation rule orders neighbours by Greville distance and resolves ties by
index order; identical abscissae give identical neighbour orders.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import cached_property

import numpy as np
from scipy.interpolate import BSpline


@dataclass(frozen=True)
class SplineBasis1D:
    """Clamped uniform B-spline basis of ``degree`` on ``n_elements`` cells."""

    degree: int
    n_elements: int
    knots: np.ndarray
    greville: np.ndarray

    @property
    def n_basis(self) -> int:
        return self.n_elements + self.degree

    @property
    def h(self) -> float:
        return 1.0 / self.n_elements

    @cached_property
    def supports(self) -> np.ndarray:
        p = self.degree
        return np.array([(self.knots[i], self.knots[i + p + 1]) for i in range(self.n_basis)])

    # -- evaluation -------------------------------------------------------
    def values(self, x: np.ndarray) -> np.ndarray:
        """Values of all basis functions, shape ``(n_basis, len(x))``."""
        return basis_value_table(self, x, nu=0)

    def derivatives(self, x: np.ndarray) -> np.ndarray:
        """First derivatives of all basis functions, shape ``(n_basis, len(x))``."""
        return basis_value_table(self, x, nu=1)


def make_open_uniform_basis(n_elements: int = 32, degree: int = 3) -> SplineBasis1D:
    """Clamped uniform basis (``code/cut_spline_sbp.py:52-76``)."""
    if n_elements < 2:
        raise ValueError("n_elements must be at least 2.")
    if degree < 1:
        raise ValueError("degree must be at least 1.")
    h = 1.0 / n_elements
    knots = np.concatenate(
        [np.zeros(degree + 1), np.arange(1, n_elements, dtype=float) * h, np.ones(degree + 1)]
    )
    n_basis = len(knots) - degree - 1
    greville = np.array([np.sum(knots[i + 1 : i + degree + 1]) / degree for i in range(n_basis)])
    return SplineBasis1D(degree=degree, n_elements=n_elements, knots=knots, greville=greville)


def basis_value_table(basis: SplineBasis1D, x: np.ndarray, nu: int = 0) -> np.ndarray:
    """All basis values (``nu=0``) or first derivatives (``nu=1``) at ``x``.

    Points are clamped to ``[0, 1]``. The derivative uses
    ``B'_{i,k} = k (B_{i,k-1}/(t_{i+k}-t_i) - B_{i+1,k-1}/(t_{i+k+1}-t_{i+1}))``
    on the same knot vector; at the right end point the lower-degree table is
    taken one ulp to the left, where the last cell's polynomial is exact.
    """
    t = basis.knots
    k = basis.degree
    n = basis.n_basis
    x = np.asarray(x, dtype=float)
    if x.size == 0:
        return np.zeros((n, 0))
    xc = np.clip(x, t[k], t[-k - 1])
    if nu == 0:
        return np.asarray(BSpline.design_matrix(xc, t, k, extrapolate=False).todense()).T
    if nu != 1:
        raise ValueError("only nu in {0, 1} is supported")
    right = t[-k - 1]
    xc = np.where(xc >= right, np.nextafter(right, t[k]), xc)
    low = np.asarray(BSpline.design_matrix(xc, t, k - 1, extrapolate=False).todense())
    da = t[k : k + n] - t[:n]
    db = t[k + 1 : k + 1 + n] - t[1 : 1 + n]
    ca = np.where(da > 0.0, k / np.where(da > 0.0, da, 1.0), 0.0)
    cb = np.where(db > 0.0, k / np.where(db > 0.0, db, 1.0), 0.0)
    d = low[:, :n] * ca[None, :] - low[:, 1 : 1 + n] * cb[None, :]
    return d.T


def element_support(basis: SplineBasis1D, element: int) -> tuple[int, int]:
    """Index range ``[e, e+p+1)`` of the basis functions alive on element ``e``."""
    return element, element + basis.degree + 1


def local_tables(basis: SplineBasis1D, elements: np.ndarray, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Values and derivatives of the ``p+1`` local functions at points ``x``.

    ``elements[j]`` is the element that owns point ``x[j]``. Returns two arrays of
    shape ``(p+1, len(x))`` whose row ``a`` is basis function ``elements[j] + a``.
    """
    p = basis.degree
    elements = np.asarray(elements, dtype=np.int64)
    V = basis_value_table(basis, x, 0)
    D = basis_value_table(basis, x, 1)
    cols = np.arange(x.size)
    rows = elements[None, :] + np.arange(p + 1)[:, None]
    return V[rows, cols[None, :]], D[rows, cols[None, :]]


def collocation_matrix(basis: SplineBasis1D) -> np.ndarray:
    """``A[j, i] = N_i(greville_j)`` evaluated one basis function at a time.

    Each function is a ``BSpline`` with a unit coefficient vector and
    ``extrapolate=False`` (NaN outside the support mapped to 0), the same
    evaluation path the reference uses for its collocation solve.
    """
    n = basis.n_basis
    eye = np.eye(n)
    rows = []
    for i in range(n):
        f = BSpline(basis.knots, eye[i], basis.degree, extrapolate=False)
        rows.append(np.nan_to_num(f(basis.greville), nan=0.0))
    return np.vstack(rows).T


def polynomial_coefficients(basis: SplineBasis1D, max_degree: int) -> np.ndarray:
    """Spline coefficients of ``x**r`` for ``r = 0..max_degree`` (rows).

    Follows ``code/cut_spline_sbp.py:148-153``: collocation at the Greville
    abscissae, one ``np.linalg.solve`` per monomial.
    """
    A = collocation_matrix(basis)
    return np.asarray([np.linalg.solve(A, basis.greville**r) for r in range(max_degree + 1)])


def gauss_legendre(order: int) -> tuple[np.ndarray, np.ndarray]:
    return np.polynomial.legendre.leggauss(order)


def interval_forms(basis: SplineBasis1D, element: int, a: float, b: float, order: int):
    """Local 1D mass and stiffness on ``[a, b]`` inside one element.

    Returns ``(lo, M, K, xq, wq, V)`` with ``M``/``K`` of shape ``(p+1, p+1)`` and
    ``V`` the local values at the Gauss points (``code/bell_3d_core.py:172-200``).
    """
    g, w = gauss_legendre(order)
    xq = 0.5 * (b - a) * g + 0.5 * (b + a)
    wq = 0.5 * (b - a) * w
    lo, hi = element_support(basis, element)
    V = basis_value_table(basis, xq, 0)[lo:hi]
    D = basis_value_table(basis, xq, 1)[lo:hi]
    return lo, (V * wq) @ V.T, (D * wq) @ D.T, xq, wq, V


def full_cell_forms(basis: SplineBasis1D, order: int) -> tuple[np.ndarray, np.ndarray]:
    """Per-element full-cell mass and stiffness, shape ``(n_elements, p+1, p+1)``."""
    p = basis.degree
    Ms = np.empty((basis.n_elements, p + 1, p + 1))
    Ks = np.empty_like(Ms)
    h = basis.h
    for e in range(basis.n_elements):
        _, M, K, _, _, _ = interval_forms(basis, e, e * h, (e + 1) * h, order)
        Ms[e] = M
        Ks[e] = K
    return Ms, Ks
