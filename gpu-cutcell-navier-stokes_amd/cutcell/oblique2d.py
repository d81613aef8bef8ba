"""2D half-plane cut-cell systems: assembly, active sets, aggregation ``E``.

This module builds the inputs of the PCG path for the 2D configurations (the
manufactured-solution oblique cut, the sliver stress sweep, and the batched
oblique sweep). Everything is assembled sparse; the reference assembles dense
``n_full x n_full`` blocks, which does not scale to a batch of 48x48 systems.

Reference behaviour followed (reference repository, file:line):
- half-plane cut ``n . x >= offset`` and ``from_angle``: ``code/oblique_cut_sbp.py:17-28``;
- Sutherland-Hodgman clip with inside tolerance 1e-14 and duplicate removal at
  1e-13: ``code/oblique_cut_sbp.py:106-147``;
- fan triangulation with product Gauss-Duffy rules and edge rules:
  ``code/oblique_cut_sbp.py:210-262``;
- per-cell Gram updates ``H, Qx, Qy, K`` and boundary forms ``Sx, Sy`` on edges
  lying on the cut or on the unit square: ``code/oblique_cut_sbp.py:271-375``;
  dof index ``iy * nx + ix`` (``:69-71``);
- active / interior / small sets from ``gamma = H_ii / H_bg_ii``:
  ``code/oblique_cut_sbp.py:575-588``;
- total-degree monomial rows ordered by (total, s_x): ``:596-611``;
- the polynomial-preserving extension (nearest interior Greville points in a
  stable distance order, smallest full-rank prefix at tol 1e-12, min-norm
  least squares): ``code/oblique_cut_sbp.py:614-672``;
- the congruence ``E^T X E``: ``code/oblique_cut_sbp.py:693-707``, and the SPD
  operator ``sym(E^T K E + E^T H E)``: ``code/fast_solver_workbench.py:159-165``.

Uncut cells are integrated with the tensor Gauss rule of the 1D forms (exact
for the spline products), cut cells with the clipped-polygon Duffy rule; both
are exact for the polynomial integrands, so values agree with the all-Duffy
reference assembly to round-off.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from math import cos, sin

import numpy as np
import scipy.sparse as sp

from .basis import (
    SplineBasis1D,
    basis_value_table,
    gauss_legendre,
    make_open_uniform_basis,
    polynomial_coefficients,
)


@dataclass(frozen=True)
class HalfPlaneCut:
    """The unit square cut by the half plane ``normal . x >= offset``."""

    normal: np.ndarray
    offset: float

    @staticmethod
    def from_angle(angle: float, offset: float) -> "HalfPlaneCut":
        n = np.array([cos(angle), sin(angle)], dtype=float)
        n = n / np.linalg.norm(n)
        return HalfPlaneCut(normal=n, offset=float(offset))


# ---------------------------------------------------------------------------
# Polygon clipping and quadrature.
# ---------------------------------------------------------------------------


def _inside(point: np.ndarray, cut: HalfPlaneCut) -> bool:
    return bool(np.dot(cut.normal, point) >= cut.offset - 1.0e-14)


def _crossing(a: np.ndarray, b: np.ndarray, cut: HalfPlaneCut) -> np.ndarray:
    fa = float(np.dot(cut.normal, a) - cut.offset)
    fb = float(np.dot(cut.normal, b) - cut.offset)
    denom = fa - fb
    if abs(denom) < 1.0e-30:
        return 0.5 * (a + b)
    t = min(1.0, max(0.0, fa / denom))
    return a + t * (b - a)


def clip_polygon_halfplane(poly: np.ndarray, cut: HalfPlaneCut) -> np.ndarray:
    """Clip a counter-clockwise polygon against ``normal . x >= offset``."""
    poly = np.asarray(poly, dtype=float)
    if len(poly) == 0:
        return np.empty((0, 2))
    out: list[np.ndarray] = []
    prev = poly[-1]
    prev_in = _inside(prev, cut)
    for cur in poly:
        cur_in = _inside(cur, cut)
        if cur_in:
            if not prev_in:
                out.append(_crossing(prev, cur, cut))
            out.append(cur.copy())
        elif prev_in:
            out.append(_crossing(prev, cur, cut))
        prev, prev_in = cur, cur_in
    kept: list[np.ndarray] = []
    for pt in out:
        if not kept or np.linalg.norm(pt - kept[-1]) > 1.0e-13:
            kept.append(pt)
    if len(kept) > 1 and np.linalg.norm(kept[0] - kept[-1]) <= 1.0e-13:
        kept.pop()
    return np.asarray(kept, dtype=float).reshape(-1, 2)


def polygon_area(poly: np.ndarray) -> float:
    if len(poly) < 3:
        return 0.0
    x, y = poly[:, 0], poly[:, 1]
    return float(0.5 * np.sum(x * np.roll(y, -1) - y * np.roll(x, -1)))


def triangle_rule(v0, v1, v2, order: int):
    """Collapsed (Duffy) product Gauss rule on a triangle, ``order**2`` points."""
    g, w = gauss_legendre(order)
    u = 0.5 * (g + 1.0)
    wu = 0.5 * w
    e1 = v1 - v0
    e2 = v2 - v0
    det = abs(float(e1[0] * e2[1] - e1[1] * e2[0]))
    U = np.repeat(u, order)
    V = np.tile(u, order)
    W = det * (1.0 - U) * np.repeat(wu, order) * np.tile(wu, order)
    pts = v0[None, :] + U[:, None] * e1[None, :] + ((1.0 - U) * V)[:, None] * e2[None, :]
    return pts[:, 0], pts[:, 1], W


def polygon_rule(poly: np.ndarray, order: int):
    if len(poly) < 3:
        return np.empty(0), np.empty(0), np.empty(0)
    if polygon_area(poly) < 0.0:
        poly = poly[::-1]
    xs, ys, ws = [], [], []
    for k in range(1, len(poly) - 1):
        x, y, w = triangle_rule(poly[0], poly[k], poly[k + 1], order)
        xs.append(x)
        ys.append(y)
        ws.append(w)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ws)


def edge_rule(a: np.ndarray, b: np.ndarray, order: int):
    g, w = gauss_legendre(order)
    t = 0.5 * (g + 1.0)
    length = float(np.linalg.norm(b - a))
    pts = a[:, None] + (b - a)[:, None] * t[None, :]
    if length <= 1.0e-15:
        normal = np.zeros(2)
    else:
        dx, dy = (b - a) / length
        normal = np.array([dy, -dx])
    return pts[0], pts[1], 0.5 * length * w, normal


# ---------------------------------------------------------------------------
# Assembly.
# ---------------------------------------------------------------------------


@dataclass
class ObliqueCutSystem2D:
    """Sparse cut-domain forms on the full background index space."""

    basis_x: SplineBasis1D
    basis_y: SplineBasis1D
    cut: HalfPlaneCut
    H: sp.csr_matrix
    K: sp.csr_matrix
    H_background_diag: np.ndarray
    load: np.ndarray | None = None
    Qx: sp.csr_matrix | None = None
    Qy: sp.csr_matrix | None = None
    Sx: sp.csr_matrix | None = None
    Sy: sp.csr_matrix | None = None
    n_cut_cells: int = 0
    min_cut_fraction: float = 1.0

    @property
    def nx(self) -> int:
        return self.basis_x.n_basis

    @property
    def ny(self) -> int:
        return self.basis_y.n_basis

    @property
    def n_full(self) -> int:
        return self.nx * self.ny


class _Coo:
    def __init__(self, n: int):
        self.n = n
        self.r: list[np.ndarray] = []
        self.c: list[np.ndarray] = []
        self.v: list[np.ndarray] = []

    def add(self, rows, cols, vals):
        self.r.append(np.asarray(rows, dtype=np.int64).ravel())
        self.c.append(np.asarray(cols, dtype=np.int64).ravel())
        self.v.append(np.asarray(vals, dtype=float).ravel())

    def tocsr(self) -> sp.csr_matrix:
        if not self.r:
            return sp.csr_matrix((self.n, self.n))
        return sp.coo_matrix(
            (np.concatenate(self.v), (np.concatenate(self.r), np.concatenate(self.c))), shape=(self.n, self.n)
        ).tocsr()


def _element_forms(basis: SplineBasis1D, order: int):
    """Per-element full-cell 1D mass, stiffness and ``Q = int N_a N_b'``."""
    p = basis.degree
    g, w = gauss_legendre(order)
    h = basis.h
    ne = basis.n_elements
    M = np.empty((ne, p + 1, p + 1))
    K = np.empty_like(M)
    Q = np.empty_like(M)
    for e in range(ne):
        a, b = e * h, (e + 1) * h
        xq = 0.5 * (b - a) * g + 0.5 * (b + a)
        wq = 0.5 * (b - a) * w
        V = basis_value_table(basis, xq, 0)[e : e + p + 1]
        D = basis_value_table(basis, xq, 1)[e : e + p + 1]
        M[e] = (V * wq) @ V.T
        K[e] = (D * wq) @ D.T
        Q[e] = (V * wq) @ D.T
    return M, K, Q


def _local_rows(ix: int, iy: int, p: int, nx: int) -> np.ndarray:
    return (np.arange(iy, iy + p + 1)[:, None] * nx + np.arange(ix, ix + p + 1)[None, :]).reshape(-1)


def assemble_oblique_cut(
    basis_x: SplineBasis1D,
    basis_y: SplineBasis1D,
    cut: HalfPlaneCut,
    quadrature_order: int | None = None,
    *,
    with_sbp: bool = False,
    source=None,
    flux=None,
) -> ObliqueCutSystem2D:
    """Assemble ``H`` and ``K`` (and optionally ``Qx, Qy, Sx, Sy`` and a load).

    ``source`` is an optional vectorised callable ``f(x, y)``; when given, the
    load vector ``int_Omega f phi_i`` is assembled with the same rules.
    ``flux`` is an optional callable returning ``(g_x, g_y)`` at boundary
    points; its normal component is added as the natural boundary term
    ``int_{dOmega} (g . n) phi_i``.
    The default cut-cell rule order is ``max(10, p+5)``
    (``code/oblique_cut_sbp.py:293``).
    """
    if basis_x.n_elements != basis_y.n_elements or basis_x.degree != basis_y.degree:
        raise ValueError("equal element counts and degrees are expected in x and y")
    p = basis_x.degree
    N = basis_x.n_elements
    q = quadrature_order or max(10, p + 5)
    nx, ny = basis_x.n_basis, basis_y.n_basis
    n = nx * ny
    h = 1.0 / N
    p1 = p + 1

    # Corner classification of every cell (inside test as in the clip).
    xs = np.arange(N + 1) * h
    corner_in = (cut.normal[0] * xs[None, :] + cut.normal[1] * xs[:, None]) >= cut.offset - 1.0e-14  # [iy, ix]
    all_in = corner_in[:-1, :-1] & corner_in[:-1, 1:] & corner_in[1:, :-1] & corner_in[1:, 1:]
    any_in = corner_in[:-1, :-1] | corner_in[:-1, 1:] | corner_in[1:, :-1] | corner_in[1:, 1:]

    gorder = max(p + 2, 8)
    Mx, Kx, Qx1 = _element_forms(basis_x, gorder)
    My, Ky, Qy1 = _element_forms(basis_y, gorder)

    tH, tK = _Coo(n), _Coo(n)
    tQx, tQy, tSx, tSy = _Coo(n), _Coo(n), _Coo(n), _Coo(n)
    load = np.zeros(n) if source is not None else None
    need_edges = with_sbp or flux is not None

    # --- uncut cells: tensor blocks, vectorised over cells --------------------
    fy, fx = np.nonzero(all_in)
    if fy.size:
        # local (a_y, a_x) x (b_y, b_x) blocks for every full cell.
        rows = (fy[:, None, None] + np.arange(p1)[None, :, None]) * nx + (fx[:, None, None] + np.arange(p1)[None, None, :])
        rows = rows.reshape(fy.size, p1 * p1)
        R = np.repeat(rows[:, :, None], p1 * p1, axis=2)
        C = np.repeat(rows[:, None, :], p1 * p1, axis=1)
        kron = lambda A, B: np.einsum("eab,ecd->eacbd", A, B).reshape(A.shape[0], p1 * p1, p1 * p1)
        tH.add(R, C, kron(My[fy], Mx[fx]))
        tK.add(R, C, kron(My[fy], Kx[fx]) + kron(Ky[fy], Mx[fx]))
        if with_sbp:
            tQx.add(R, C, kron(My[fy], Qx1[fx]))
            tQy.add(R, C, kron(Qy1[fy], Mx[fx]))
        if source is not None:
            g, w = gauss_legendre(gorder)
            for k in range(fy.size):
                iy, ix = int(fy[k]), int(fx[k])
                xq = 0.5 * h * g + (ix + 0.5) * h
                yq = 0.5 * h * g + (iy + 0.5) * h
                X, Y = np.meshgrid(xq, yq, indexing="xy")
                W = np.outer(0.5 * h * w, 0.5 * h * w)
                vx = basis_value_table(basis_x, xq, 0)[ix : ix + p1]
                vy = basis_value_table(basis_y, yq, 0)[iy : iy + p1]
                F = source(X, Y) * W  # [qy, qx]
                load[rows[k]] += np.einsum("aj,ji,bi->ab", vy, F, vx).reshape(-1)

    # --- cut cells: clipped-polygon Duffy rules -------------------------------
    def on_boundary(a, b) -> bool:
        mid = 0.5 * (a + b)
        on_cut = abs(float(np.dot(cut.normal, mid) - cut.offset)) <= 1.0e-10
        on_square = (
            abs(mid[0]) <= 1.0e-12 or abs(mid[0] - 1.0) <= 1.0e-12 or abs(mid[1]) <= 1.0e-12 or abs(mid[1] - 1.0) <= 1.0e-12
        )
        return bool(on_cut or on_square)

    def local_tensor(ix, iy, xq, yq):
        vx = basis_value_table(basis_x, xq, 0)[ix : ix + p1]
        dvx = basis_value_table(basis_x, xq, 1)[ix : ix + p1]
        vy = basis_value_table(basis_y, yq, 0)[iy : iy + p1]
        dvy = basis_value_table(basis_y, yq, 1)[iy : iy + p1]
        phi = (vy[:, None, :] * vx[None, :, :]).reshape(p1 * p1, xq.size)
        dpx = (vy[:, None, :] * dvx[None, :, :]).reshape(p1 * p1, xq.size)
        dpy = (dvy[:, None, :] * vx[None, :, :]).reshape(p1 * p1, xq.size)
        return phi, dpx, dpy

    n_cut = 0
    min_frac = 1.0
    cy, cx = np.nonzero(any_in & ~all_in)
    boundary_cells = set(zip(cy.tolist(), cx.tolist()))
    if need_edges:
        # Uncut cells on the square boundary carry boundary forms too.
        for iy in range(N):
            for ix in range(N):
                if all_in[iy, ix] and (ix in (0, N - 1) or iy in (0, N - 1)):
                    boundary_cells.add((iy, ix))
    for iy, ix in sorted(boundary_cells):
        x0, x1, y0, y1 = ix * h, (ix + 1) * h, iy * h, (iy + 1) * h
        cell = np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]])
        is_cut = not all_in[iy, ix]
        clipped = clip_polygon_halfplane(cell, cut) if is_cut else cell
        area = polygon_area(clipped)
        if len(clipped) < 3 or abs(area) <= 1.0e-16:
            continue
        if area < 0.0:
            clipped = clipped[::-1]
        rows = _local_rows(ix, iy, p, nx)
        R = np.repeat(rows, rows.size)
        C = np.tile(rows, rows.size)
        if is_cut:
            n_cut += 1
            min_frac = min(min_frac, abs(area) / (h * h))
            xq, yq, wq = polygon_rule(clipped, q)
            phi, dpx, dpy = local_tensor(ix, iy, xq, yq)
            tH.add(R, C, (phi * wq) @ phi.T)
            tK.add(R, C, (dpx * wq) @ dpx.T + (dpy * wq) @ dpy.T)
            if with_sbp:
                tQx.add(R, C, (phi * wq) @ dpx.T)
                tQy.add(R, C, (phi * wq) @ dpy.T)
            if source is not None:
                load[rows] += phi @ (source(xq, yq) * wq)
        if need_edges:
            for k in range(len(clipped)):
                a = clipped[k]
                b = clipped[(k + 1) % len(clipped)]
                if not on_boundary(a, b):
                    continue
                xe, ye, we, nrm = edge_rule(a, b, q)
                if np.allclose(nrm, 0.0):
                    continue
                phe, _, _ = local_tensor(ix, iy, xe, ye)
                if with_sbp:
                    bm = (phe * we) @ phe.T
                    tSx.add(R, C, nrm[0] * bm)
                    tSy.add(R, C, nrm[1] * bm)
                if flux is not None and load is not None:
                    gx, gy = flux(xe, ye)
                    load[rows] += phe @ ((gx * nrm[0] + gy * nrm[1]) * we)

    # Background mass diagonal: kron of the 1D full-interval masses.
    m1x = np.zeros(nx)
    m1y = np.zeros(ny)
    for e in range(N):
        m1x[e : e + p1] += np.diag(Mx[e])
        m1y[e : e + p1] += np.diag(My[e])
    hbg = np.kron(m1y, m1x)

    return ObliqueCutSystem2D(
        basis_x=basis_x,
        basis_y=basis_y,
        cut=cut,
        H=tH.tocsr(),
        K=tK.tocsr(),
        H_background_diag=hbg,
        load=load,
        Qx=tQx.tocsr() if with_sbp else None,
        Qy=tQy.tocsr() if with_sbp else None,
        Sx=tSx.tocsr() if with_sbp else None,
        Sy=tSy.tocsr() if with_sbp else None,
        n_cut_cells=n_cut,
        min_cut_fraction=min_frac,
    )


# ---------------------------------------------------------------------------
# Active sets and the polynomial-preserving extension E.
# ---------------------------------------------------------------------------


@dataclass
class Aggregation2D:
    E: sp.csr_matrix
    interior: np.ndarray
    small_support: np.ndarray
    active: np.ndarray
    gamma: np.ndarray
    polynomial_degree: int
    polynomial_residual: float
    stencils: dict = field(default_factory=dict)


def active_sets(system: ObliqueCutSystem2D, support_threshold: float = 0.05, active_tol: float = 1.0e-22):
    """``gamma = H_ii / H_bg_ii``; active ``gamma > tol``; interior ``gamma >= thr``."""
    bg = system.H_background_diag
    phys = system.H.diagonal()
    gamma = np.zeros_like(phys)
    pos = bg > 0.0
    gamma[pos] = phys[pos] / bg[pos]
    active = np.nonzero(gamma > active_tol)[0]
    interior = np.nonzero(gamma >= support_threshold)[0]
    small = active[gamma[active] < support_threshold]
    return active, interior, small.astype(np.int64), gamma


def tensor_greville(basis_x: SplineBasis1D, basis_y: SplineBasis1D) -> np.ndarray:
    xx, yy = np.meshgrid(basis_x.greville, basis_y.greville, indexing="xy")
    return np.column_stack([xx.ravel(), yy.ravel()])


def tensor_polynomial_rows(basis_x: SplineBasis1D, basis_y: SplineBasis1D, degree: int) -> np.ndarray:
    """Rows ``kron(cy[s_y], cx[s_x])`` for ``s_x + s_y <= degree``, by (total, s_x)."""
    cx = polynomial_coefficients(basis_x, degree)
    cy = polynomial_coefficients(basis_y, degree)
    rows = []
    for total in range(degree + 1):
        for sx in range(total + 1):
            rows.append(np.kron(cy[total - sx], cx[sx]))
    return np.asarray(rows)


def polynomial_preserving_extension(
    system: ObliqueCutSystem2D,
    support_threshold: float = 0.05,
    preserve_degree: int | None = None,
    active_tol: float = 1.0e-22,
) -> Aggregation2D:
    """Build the sparse extension ``E`` (``n_full x n_interior``)."""
    bx, by = system.basis_x, system.basis_y
    active, interior, small, gamma = active_sets(system, support_threshold, active_tol)
    if interior.size == 0:
        raise ValueError("No interior active tensor-product basis functions were found.")
    requested = bx.degree if preserve_degree is None else preserve_degree
    by_count = 0
    while (by_count + 1) * (by_count + 2) // 2 <= interior.size:
        by_count += 1
    degree = min(requested, by_count - 1)
    poly = tensor_polynomial_rows(bx, by, degree)
    n_terms = poly.shape[0]
    n_full = system.n_full
    n_red = interior.size
    col = np.full(n_full, -1, dtype=np.int64)
    col[interior] = np.arange(n_red)

    rows_t = [interior.astype(np.int64)]
    cols_t = [np.arange(n_red, dtype=np.int64)]
    vals_t = [np.ones(n_red)]
    greville = tensor_greville(bx, by)
    g_int = greville[interior]
    stencils = {}
    for b in small:
        dist = np.linalg.norm(g_int - greville[int(b)], axis=1)
        ordered = interior[np.argsort(dist, kind="stable")]
        chosen = ordered[:n_terms]
        for k in range(n_terms, ordered.size + 1):
            cand = ordered[:k]
            if np.linalg.matrix_rank(poly[:, cand], tol=1.0e-12) == n_terms:
                chosen = cand
                break
        w = np.linalg.lstsq(poly[:, chosen], poly[:, int(b)], rcond=None)[0]
        stencils[int(b)] = (chosen.copy(), w.copy())
        rows_t.append(np.full(chosen.size, int(b), dtype=np.int64))
        cols_t.append(col[chosen])
        vals_t.append(w)
    E = sp.coo_matrix(
        (np.concatenate(vals_t), (np.concatenate(rows_t), np.concatenate(cols_t))), shape=(n_full, n_red)
    ).tocsr()
    E.eliminate_zeros()
    active_mask = gamma > active_tol
    res = 0.0
    for row in poly:
        res = max(res, float(np.max(np.abs(E @ row[interior] - row * active_mask))))
    return Aggregation2D(
        E=E,
        interior=interior.astype(np.int64),
        small_support=small,
        active=active.astype(np.int64),
        gamma=gamma,
        polynomial_degree=degree,
        polynomial_residual=res,
        stencils=stencils,
    )


def pullback(X: sp.spmatrix, E: sp.csr_matrix) -> sp.csr_matrix:
    """``E^T X E`` (``code/oblique_cut_sbp.py:693-707``)."""
    return (E.T @ X @ E).tocsr()


def sym(A: sp.spmatrix) -> sp.csr_matrix:
    B = (0.5 * (A + A.T)).tocsr()
    B.eliminate_zeros()
    B.sort_indices()
    return B


# ---------------------------------------------------------------------------
# System recipes.
# ---------------------------------------------------------------------------


@dataclass
class CutSystem:
    """An aggregated SPD system ``A x = b`` with its provenance."""

    A: sp.csr_matrix
    b: np.ndarray
    E: sp.csr_matrix
    interior: np.ndarray
    small_support: np.ndarray
    active: np.ndarray
    meta: dict


def sliver_system(n_elements: int, degree: int, epsilon: float) -> CutSystem:
    """The sliver stress case (``code/fast_solver_workbench.py:138-168``).

    Vertical cut at ``x = h(1 - eps)``, rule order ``max(24, 2p+14)``,
    ``A = sym(E^T K E + E^T H E)``, Gaussian RHS with seed ``12345 + 17 N + p``.
    """
    bx = make_open_uniform_basis(n_elements, degree)
    by = make_open_uniform_basis(n_elements, degree)
    cut = HalfPlaneCut.from_angle(0.0, bx.h * (1.0 - epsilon))
    sysm = assemble_oblique_cut(bx, by, cut, quadrature_order=max(24, 2 * degree + 14))
    agg = polynomial_preserving_extension(sysm, 0.05, degree, active_tol=1.0e-100)
    A = sym(pullback(sysm.K, agg.E) + pullback(sysm.H, agg.E))
    rng = np.random.default_rng(12345 + 17 * n_elements + degree)
    b = rng.standard_normal(A.shape[0])
    meta = dict(kind="sliver", n_elements=n_elements, degree=degree, epsilon=epsilon, n_full=sysm.n_full)
    return CutSystem(A, b, agg.E, agg.interior, agg.small_support, agg.active, meta)


def oblique_system(
    n_elements: int,
    degree: int,
    angle: float,
    offset: float,
    *,
    quadrature_order: int | None = None,
    rhs: str = "manufactured",
    seed: int = 0,
    active_tol: float = 1.0e-22,
) -> CutSystem:
    """Screened Poisson ``-lap u + u = f`` on the half-plane-cut unit square.

    ``rhs="manufactured"`` uses ``f = (2 pi^2 + 1) sin(pi x) sin(pi y)`` and the
    natural boundary data ``grad u . n`` of ``u = sin(pi x) sin(pi y)`` on the
    cut and on the square, so ``u`` solves the weak problem the matrix
    discretises; ``rhs="gaussian"`` draws ``N(0, 1)`` entries from
    ``default_rng(seed)``.
    """
    bx = make_open_uniform_basis(n_elements, degree)
    by = make_open_uniform_basis(n_elements, degree)
    cut = HalfPlaneCut.from_angle(angle, offset)
    q = quadrature_order or max(24, 2 * degree + 14)
    f = g = None
    if rhs == "manufactured":
        f = lambda x, y: (2.0 * np.pi**2 + 1.0) * np.sin(np.pi * x) * np.sin(np.pi * y)
        g = lambda x, y: (np.pi * np.cos(np.pi * x) * np.sin(np.pi * y), np.pi * np.sin(np.pi * x) * np.cos(np.pi * y))
    sysm = assemble_oblique_cut(bx, by, cut, quadrature_order=q, source=f, flux=g)
    agg = polynomial_preserving_extension(sysm, 0.05, degree, active_tol=active_tol)
    A = sym(pullback(sysm.K, agg.E) + pullback(sysm.H, agg.E))
    if rhs == "manufactured":
        b = np.asarray(agg.E.T @ sysm.load)
    else:
        b = np.random.default_rng(seed).standard_normal(A.shape[0])
    meta = dict(
        kind="oblique",
        n_elements=n_elements,
        degree=degree,
        angle=angle,
        offset=offset,
        n_full=sysm.n_full,
        n_cut_cells=sysm.n_cut_cells,
        min_cut_fraction=sysm.min_cut_fraction,
        polynomial_residual=agg.polynomial_residual,
    )
    return CutSystem(A, b, agg.E, agg.interior, agg.small_support, agg.active, meta)
