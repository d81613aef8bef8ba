"""The hollow-cube "bell" 3D system: the reference's published GPU benchmark input.

Builds ``A = sym(E^T K E) + sym(E^T H E)`` and ``b = E^T load`` on the bell
geometry, the system behind the reference's GPU-versus-CPU table
(``benchmarks/gpu_pcg_benchmark_3d.py:49-57``; numbers in
``benchmarks/results/gpu_solver_benchmark_3d.csv``).

Reference behaviour followed (reference repository, file:line):
- geometry: nine axis-aligned solid boxes anchored to a 20-cell grid and
  shifted by ``epsilon * h``: ``code/bell_3d_core.py:66-103``; fluid test of a
  sub-box midpoint with tolerance 1e-14: ``:49-54, 114-119``;
- each background cell splits at box edges into sub-boxes; fluid sub-boxes
  get Kronecker products of exact 1D Gauss forms (order ``max(5, p+3)``):
  ``code/bell_3d_core.py:128-144, 172-200, 338-390``;
- the Gaussian load at (0.1, 0.5, 0.5), sigma 0.06: ``:270-273, 370-373``;
- symmetrisation ``(H + H^T)/2`` and ``gamma`` from the full-cell mass:
  ``:392-406``;
- the 3D extension and ``restrict_spd``: ``:471-552`` (via ``cutcell.cut3d``).

Assembly is vectorised: for every combination of sub-interval slots the
fluid mask times the separable 1D forms is added to the stencil slab of the
corresponding dof shift, which gives the same sums as the reference's
per-cell loop up to the order of floating-point additions.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .basis import gauss_legendre, make_open_uniform_basis
from .cut3d import CutAssembly3D, Stencil3D, local_values, polynomial_preserving_extension_3d, restrict_spd


@dataclass(frozen=True)
class Box:
    x0: float
    x1: float
    y0: float
    y1: float
    z0: float
    z1: float


def bell_boxes(n_geometry: int, epsilon: float) -> tuple[Box, ...]:
    """The nine solid boxes (``code/bell_3d_core.py:66-103``)."""
    if not (0.0 < epsilon < 0.35):
        raise ValueError("epsilon must be in (0, 0.35)")
    n = int(n_geometry)
    if n < 20:
        raise ValueError("the 3D bell needs n_geometry >= 20")
    h = 1.0 / n
    s = epsilon * h
    a = round(0.2 * n)
    w = round(0.2 * n)
    b = n - a
    c0 = a + w
    c1 = b - w
    m0 = round(0.45 * n)
    m1 = n - m0

    def bx(x0, x1, y0, y1, z0, z1):
        return Box(x0 * h + s, x1 * h + s, y0 * h + s, y1 * h + s, z0 * h + s, z1 * h + s)

    return (
        bx(a, c0, a, m0, a, b),
        bx(a, c0, m1, b, a, b),
        bx(a, c0, m0, m1, a, m0),
        bx(a, c0, m0, m1, m1, b),
        bx(c1, b, a, b, a, b),
        bx(c0, c1, a, c0, a, b),
        bx(c0, c1, c1, b, a, b),
        bx(c0, c1, c0, c1, a, c0),
        bx(c0, c1, c0, c1, c1, b),
    )


def _is_solid(boxes, x, y, z, tol=1.0e-14):
    solid = np.zeros(np.broadcast(x, y, z).shape, dtype=bool)
    for bx in boxes:
        solid |= ((bx.x0 - tol <= x) & (x <= bx.x1 + tol) & (bx.y0 - tol <= y) & (y <= bx.y1 + tol)
                  & (bx.z0 - tol <= z) & (z <= bx.z1 + tol))
    return solid


def _axis_pieces(basis, edges, order, centre, sigma):
    """Split every element at the box edges inside it.

    Returns a list over slots s of dicts with arrays over elements:
    ``present`` (bool), ``mid`` (midpoint), ``M``, ``K`` ((p+1, p+1) forms) and
    ``L`` ((p+1) Gaussian load moments), for the s-th sub-interval of each
    element.
    """
    N = basis.n_elements
    p = basis.degree
    h = 1.0 / N
    g, w = gauss_legendre(order)
    per_elem = []
    for e in range(N):
        lo, hi = e * h, (e + 1) * h
        cuts = sorted(set([lo, hi] + [v for v in edges if lo < v < hi]))
        per_elem.append([(a, b) for a, b in zip(cuts[:-1], cuts[1:]) if b > a])
    smax = max(len(x) for x in per_elem)
    slots = []
    for s in range(smax):
        present = np.zeros(N, dtype=bool)
        mid = np.full(N, np.nan)
        M = np.zeros((N, p + 1, p + 1))
        K = np.zeros((N, p + 1, p + 1))
        L = np.zeros((N, p + 1))
        for e in range(N):
            if s >= len(per_elem[e]):
                continue
            a, b = per_elem[e][s]
            xq = 0.5 * (b - a) * g + 0.5 * (b + a)
            wq = 0.5 * (b - a) * w
            V, D = local_values(basis, np.full(order, e), xq)
            present[e] = True
            mid[e] = 0.5 * (a + b)
            M[e] = (V * wq) @ V.T
            K[e] = (D * wq) @ D.T
            L[e] = (V * wq) @ np.exp(-((xq - centre) ** 2) / (2.0 * sigma**2))
        slots.append(dict(present=present, mid=mid, M=M, K=K, L=L))
    return slots


@dataclass
class BellSystem:
    A: object
    b: np.ndarray
    E: object
    meta: dict


def assemble_bell(n_elements: int, degree: int = 2, epsilon: float = 1.0e-3, n_geometry: int = 20,
                  pulse_center=(0.10, 0.50, 0.50), pulse_sigma: float = 0.06):
    N = n_elements
    p = degree
    p1 = p + 1
    q = max(5, p + 3)
    boxes = bell_boxes(n_geometry, epsilon)
    basis = make_open_uniform_basis(N, p)
    nb = basis.n_basis
    edges = [sorted({v for bx in boxes for v in ((bx.x0, bx.x1), (bx.y0, bx.y1), (bx.z0, bx.z1))[ax]}) for ax in range(3)]
    pieces = [_axis_pieces(basis, edges[ax], q, pulse_center[ax], pulse_sigma) for ax in range(3)]
    H = Stencil3D(nb, nb, nb, p)
    K = Stencil3D(nb, nb, nb, p)
    load = np.zeros((nb, nb, nb))
    Hv = H.data.reshape(-1, nb, nb, nb)
    Kv = K.data.reshape(-1, nb, nb, nb)
    frac = np.zeros((N, N, N))
    for sz in pieces[2]:
        for sy in pieces[1]:
            for sx in pieces[0]:
                pres = sz["present"][:, None, None] & sy["present"][None, :, None] & sx["present"][None, None, :]
                if not pres.any():
                    continue
                Z, Y, X = np.meshgrid(sz["mid"], sy["mid"], sx["mid"], indexing="ij")
                inside = (X >= 0) & (X <= 1) & (Y >= 0) & (Y <= 1) & (Z >= 0) & (Z <= 1)
                mask = (pres & inside & ~_is_solid(boxes, X, Y, Z)).astype(float)
                if not mask.any():
                    continue
                for az in range(p1):
                    for bz in range(p1):
                        mz = sz["M"][:, az, bz][:, None, None]
                        kz = sz["K"][:, az, bz][:, None, None]
                        for ay in range(p1):
                            for by in range(p1):
                                my = sy["M"][:, ay, by][None, :, None]
                                ky = sy["K"][:, ay, by][None, :, None]
                                mzy = mask * (mz * my)
                                kzy = mask * (kz * my + mz * ky)
                                for ax in range(p1):
                                    for bx_ in range(p1):
                                        o = H.offset_index(bz - az, by - ay, bx_ - ax)
                                        mx = sx["M"][:, ax, bx_][None, None, :]
                                        kx = sx["K"][:, ax, bx_][None, None, :]
                                        Hv[o, az : az + N, ay : ay + N, ax : ax + N] += mzy * mx
                                        Kv[o, az : az + N, ay : ay + N, ax : ax + N] += kzy * mx + mzy * kx
                for az in range(p1):
                    for ay in range(p1):
                        for ax in range(p1):
                            load[az : az + N, ay : ay + N, ax : ax + N] += mask * (
                                sz["L"][:, az][:, None, None] * sy["L"][:, ay][None, :, None] * sx["L"][:, ax][None, None, :]
                            )
    Hc = H.tocsr()
    Kc = K.tocsr()
    Hc = (0.5 * (Hc + Hc.T)).tocsr()
    Kc = (0.5 * (Kc + Kc.T)).tocsr()
    # Background (full-cell) mass diagonal.
    m1 = np.zeros(nb)
    g, w = gauss_legendre(q)
    h = 1.0 / N
    for e in range(N):
        xq = 0.5 * h * g + (e + 0.5) * h
        V, _ = local_values(basis, np.full(q, e), xq)
        m1[e : e + p1] += np.sum(V * V * (0.5 * h * w), axis=1)
    hbg = np.einsum("k,j,i->kji", m1, m1, m1).reshape(-1)
    diag = Hc.diagonal()
    gamma = np.zeros(nb**3)
    pos = hbg > 0
    gamma[pos] = diag[pos] / hbg[pos]
    active = np.nonzero(gamma > 0.0)[0]
    asm = CutAssembly3D(bases=(basis, basis, basis), H=Hc, K=Kc, h_background_diag=hbg, gamma=gamma, active=active,
                        cell_fraction=frac, n_cut_cells=0)
    return asm, load.reshape(-1)


def bell_system(n_elements: int, degree: int = 2, epsilon: float = 1.0e-3, tie_order: str = "stable") -> BellSystem:
    """``A = sym(E^T K E) + sym(E^T H E)``, ``b = E^T load`` on the bell."""
    asm, load = assemble_bell(n_elements, degree, epsilon)
    agg = polynomial_preserving_extension_3d(asm, tie_order=tie_order)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    b = np.asarray(agg.E.T @ load)
    meta = dict(kind="bell3d", n_elements=n_elements, degree=degree, epsilon=epsilon, n_unknowns=int(A.shape[0]),
                nnz=int(A.nnz), n_small=int(agg.small_support.size), tie_order=tie_order)
    return BellSystem(A=A, b=b, E=agg.E, meta=meta)
