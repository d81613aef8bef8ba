"""3D cut-cell systems on tensor B-spline grids: assembly, aggregation, restriction.

Builds the inputs of the PCG path for the 3D configurations (planar cut,
sphere obstacle, wall bump). The reference has no 3D code for these cut
shapes; what it does have, and what is followed here (reference repository,
file:line):
- dof order x fastest, ``(z, y, x)``: ``code/bell_3d_core.py:313-317``;
- background mass diagonal from full-cell forms and ``gamma = H_ii/H_bg_ii``,
  active ``gamma > 0``: ``code/bell_3d_core.py:343-354, 401-406``;
- the 3D polynomial-preserving extension (interior ``gamma >= 0.04``;
  monomials by (total, a, b); neighbours by Greville distance; a first stencil
  of ``n_terms + 8`` grown by 8 while ``matrix_rank(C, 1e-10) < n_terms`` and
  the stencil is shorter than ``min(#interior, 6 n_terms)``; least squares;
  weights with ``|w| <= 1e-14`` dropped): ``code/bell_3d_core.py:471-542``;
- ``restrict_spd(A, E) = sym(E^T A E)`` and ``A = restrict(K) + restrict(H)``:
  ``code/bell_3d_core.py:550-552``, ``benchmarks/gpu_pcg_benchmark_3d.py:49-57``.

Geometry is a level set ``phi >= 0`` (fluid). Each background cell is split
into the six Kuhn tetrahedra of its main diagonal; ``phi`` is interpolated
linearly on each tetrahedron, which gives a piecewise-planar fluid boundary
(exact for a planar cut). Clipped tetrahedra are decomposed into tetrahedra
and integrated with a collapsed (Duffy) Gauss product rule. Uncut cells use the
exact tensor Gauss rule. Neighbour ties in the extension are broken by dof
index (a stable sort): the reference uses NumPy's default unstable argsort,
whose tie order depends on the host's SIMD sort.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np
import scipy.sparse as sp

from .basis import SplineBasis1D, gauss_legendre, make_open_uniform_basis, polynomial_coefficients

# ---------------------------------------------------------------------------
# Local B-spline evaluation (Cox-de Boor on one element).
# ---------------------------------------------------------------------------


def _local_basis_degree(knots: np.ndarray, span: np.ndarray, x: np.ndarray, p: int) -> np.ndarray:
    """Nonzero degree-``p`` basis values on knot span ``span`` (per point).

    Returns ``(p+1, npts)``; row ``a`` is basis function ``span - p + a``.
    """
    npts = x.size
    N = np.zeros((p + 1, npts))
    N[0] = 1.0
    left = np.zeros((p + 1, npts))
    right = np.zeros((p + 1, npts))
    for j in range(1, p + 1):
        left[j] = x - knots[span + 1 - j]
        right[j] = knots[span + j] - x
        saved = np.zeros(npts)
        for r in range(j):
            den = right[r + 1] + left[j - r]
            tmp = np.where(den != 0.0, N[r] / np.where(den != 0.0, den, 1.0), 0.0)
            N[r] = saved + right[r + 1] * tmp
            saved = left[j - r] * tmp
        N[j] = saved
    return N


def local_values(basis: SplineBasis1D, elements: np.ndarray, x: np.ndarray):
    """Values and first derivatives of the ``p+1`` functions alive on each point's element.

    ``elements[j]`` owns ``x[j]``; row ``a`` of the outputs is function
    ``elements[j] + a`` (clamped open-uniform knots: span = element + p).
    """
    p = basis.degree
    t = basis.knots
    e = np.asarray(elements, dtype=np.int64)
    x = np.asarray(x, dtype=float)
    span = e + p
    V = _local_basis_degree(t, span, x, p)
    # Degree-lowering derivative: N'_{i,p} = p (N_{i,p-1}/(t_{i+p}-t_i) - N_{i+1,p-1}/(t_{i+p+1}-t_{i+1})).
    L = _local_basis_degree(t, span, x, p - 1)  # functions span-p+1 .. span
    D = np.zeros_like(V)
    for a in range(p + 1):
        i = span - p + a
        if a >= 1:
            den = t[i + p] - t[i]
            D[a] += np.where(den > 0, p * L[a - 1] / np.where(den > 0, den, 1.0), 0.0)
        if a <= p - 1:
            den = t[i + p + 1] - t[i + 1]
            D[a] -= np.where(den > 0, p * L[a] / np.where(den > 0, den, 1.0), 0.0)
    return V, D


def element_forms(basis: SplineBasis1D, order: int):
    """Per-element full-cell 1D mass and stiffness ``(n_elements, p+1, p+1)``."""
    p = basis.degree
    g, w = gauss_legendre(order)
    ne = basis.n_elements
    h = basis.h
    e = np.repeat(np.arange(ne), order)
    x = (np.arange(ne)[:, None] + 0.5 + 0.5 * g[None, :]).ravel() * h
    wq = np.tile(0.5 * h * w, ne)
    V, D = local_values(basis, e, x)
    V = V.reshape(p + 1, ne, order)
    D = D.reshape(p + 1, ne, order)
    W = wq.reshape(ne, order)
    M = np.einsum("aeq,eq,beq->eab", V, W, V)
    K = np.einsum("aeq,eq,beq->eab", D, W, D)
    return M, K


# ---------------------------------------------------------------------------
# Geometry.
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class PlanarCut:
    """Fluid is ``normal . x >= offset`` inside the unit box."""

    normal: np.ndarray
    offset: float

    def __call__(self, x, y, z):
        return self.normal[0] * x + self.normal[1] * y + self.normal[2] * z - self.offset

    @staticmethod
    def random(seed: int) -> "PlanarCut":
        """Normal uniform on the sphere; offset uniform in the middle half of
        the range of ``normal . x`` over the unit box's corners."""
        rng = np.random.default_rng(seed)
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=float)
        vals = corners @ n
        lo, hi = float(vals.min()), float(vals.max())
        c = rng.uniform(lo + 0.25 * (hi - lo), lo + 0.75 * (hi - lo))
        return PlanarCut(normal=n, offset=float(c))


@dataclass(frozen=True)
class SphereObstacle:
    """Fluid is outside the ball ``|x - center| >= radius``."""

    center: np.ndarray
    radius: float

    def __call__(self, x, y, z):
        c = self.center
        return np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - self.radius

    @staticmethod
    def random(seed: int) -> "SphereObstacle":
        """Radius uniform in [0.2, 0.4]; centre uniform such that the ball stays inside."""
        rng = np.random.default_rng(seed)
        r = rng.uniform(0.2, 0.4)
        c = rng.uniform(r, 1.0 - r, size=3)
        return SphereObstacle(center=c, radius=float(r))


# Kuhn triangulation of the unit cube: corner index = i + 2 j + 4 k.
_KUHN = []
for _perm in permutations(range(3)):
    _v = [0]
    _acc = 0
    for _ax in _perm:
        _acc |= 1 << _ax
        _v.append(_acc)
    _KUHN.append(_v)
_KUHN = np.array(_KUHN, dtype=np.int64)  # (6, 4)
_CORNERS = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=float)


def _clip_tets(P: np.ndarray, F: np.ndarray):
    """Clip tetrahedra (``P`` (T,4,3), ``F`` (T,4) level-set values) to ``F >= 0``.

    Returns sub-tetrahedra ``(S, 4, 3)`` and the index of their parent.
    """
    inside = F >= 0.0
    cnt = inside.sum(axis=1)
    out_v: list[np.ndarray] = []
    out_p: list[np.ndarray] = []
    idx_all = np.arange(P.shape[0])

    def edge(ti, a, b):
        fa = F[ti, a]
        fb = F[ti, b]
        t = fa / (fa - fb)
        return P[ti, a] + t[:, None] * (P[ti, b] - P[ti, a])

    def prism(a0, a1, a2, b0, b1, b2, parents):
        out_v.append(np.stack([a0, a1, a2, b2], axis=1))
        out_v.append(np.stack([a0, a1, b1, b2], axis=1))
        out_v.append(np.stack([a0, b0, b1, b2], axis=1))
        out_p.extend([parents, parents, parents])

    full = idx_all[cnt == 4]
    if full.size:
        out_v.append(P[full])
        out_p.append(full)
    for m in range(1, 15):
        bits = [(m >> i) & 1 for i in range(4)]
        sel = idx_all[(cnt > 0) & (cnt < 4)]
        if sel.size == 0:
            break
        code = (inside[sel] * np.array([1, 2, 4, 8])).sum(axis=1)
        ti = sel[code == m]
        if ti.size == 0:
            continue
        ins = [i for i in range(4) if bits[i]]
        outs = [i for i in range(4) if not bits[i]]
        if len(ins) == 1:
            i = ins[0]
            out_v.append(np.stack([P[ti, i], edge(ti, i, outs[0]), edge(ti, i, outs[1]), edge(ti, i, outs[2])], axis=1))
            out_p.append(ti)
        elif len(ins) == 3:
            o = outs[0]
            i0, i1, i2 = ins
            prism(P[ti, i0], P[ti, i1], P[ti, i2], edge(ti, i0, o), edge(ti, i1, o), edge(ti, i2, o), ti)
        else:
            i0, i1 = ins
            o0, o1 = outs
            prism(P[ti, i0], edge(ti, i0, o0), edge(ti, i0, o1), P[ti, i1], edge(ti, i1, o0), edge(ti, i1, o1), ti)
    if not out_v:
        return np.zeros((0, 4, 3)), np.zeros(0, dtype=np.int64)
    return np.concatenate(out_v), np.concatenate(out_p)


def _tet_rule(order: int):
    """Collapsed Gauss product rule on the reference tetrahedron (barycentric maps)."""
    g, w = gauss_legendre(order)
    u = 0.5 * (g + 1.0)
    wu = 0.5 * w
    U, V, W = np.meshgrid(u, u, u, indexing="ij")
    WU, WV, WW = np.meshgrid(wu, wu, wu, indexing="ij")
    U, V, W = U.ravel(), V.ravel(), W.ravel()
    xi1 = U
    xi2 = V * (1.0 - U)
    xi3 = W * (1.0 - U) * (1.0 - V)
    jac = (1.0 - U) ** 2 * (1.0 - V)
    weights = (WU * WV * WW).ravel() * jac
    return np.stack([xi1, xi2, xi3], axis=1), weights


# ---------------------------------------------------------------------------
# Stencil accumulation: an (n_full, (2p+1)^3) array indexed by dof and offset.
# ---------------------------------------------------------------------------


class Stencil3D:
    def __init__(self, nx: int, ny: int, nz: int, p: int):
        self.nx, self.ny, self.nz, self.p = nx, ny, nz, p
        self.w = 2 * p + 1
        self.n = nx * ny * nz
        # Offset-major: data[o, dof], so the slab of one offset is contiguous.
        self.data = np.zeros((self.w**3, self.n))

    def offset_index(self, dz, dy, dx):
        p, w = self.p, self.w
        return ((dz + p) * w + (dy + p)) * w + (dx + p)

    def tocsr(self) -> sp.csr_matrix:
        p, w = self.p, self.w
        nx, ny, nz = self.nx, self.ny, self.nz
        rows_z, rows_y, rows_x = np.unravel_index(np.arange(self.n), (nz, ny, nx))
        d = np.arange(-p, p + 1)
        DZ, DY, DX = np.meshgrid(d, d, d, indexing="ij")
        DZ, DY, DX = DZ.ravel(), DY.ravel(), DX.ravel()
        nzmask = self.data != 0.0
        o, r = np.nonzero(nzmask)
        cz = rows_z[r] + DZ[o]
        cy = rows_y[r] + DY[o]
        cx = rows_x[r] + DX[o]
        cols = (cz * ny + cy) * nx + cx
        A = sp.csr_matrix((self.data[o, r], (r, cols)), shape=(self.n, self.n))
        A.sort_indices()
        return A


# ---------------------------------------------------------------------------
# Assembly.
# ---------------------------------------------------------------------------


@dataclass
class CutAssembly3D:
    bases: tuple
    H: sp.csr_matrix
    K: sp.csr_matrix
    h_background_diag: np.ndarray
    gamma: np.ndarray
    active: np.ndarray
    cell_fraction: np.ndarray  # (nz, ny, nx) fluid volume fraction per cell
    n_cut_cells: int

    @property
    def n_full(self) -> int:
        return int(np.prod([b.n_basis for b in self.bases]))

    @property
    def min_cut_fraction(self) -> float:
        f = self.cell_fraction
        c = f[(f > 0.0) & (f < 1.0 - 1e-12)]
        return float(c.min()) if c.size else 1.0


def assemble_levelset_3d(levelset, n_elements: int, degree: int, quadrature_order: int | None = None,
                         chunk_cells: int = 96) -> CutAssembly3D:
    """Assemble fluid mass ``H`` and stiffness ``K`` on ``[0,1]^3``.

    ``levelset(x, y, z)`` is vectorised; the fluid is ``levelset >= 0``.
    """
    N = n_elements
    p = degree
    p1 = p + 1
    q = quadrature_order or (p + 3)
    bases = tuple(make_open_uniform_basis(N, p) for _ in range(3))
    nb = bases[0].n_basis
    h = 1.0 / N
    Ms, Ks = element_forms(bases[0], max(p + 2, 5))

    # Corner values and cell classes.
    xs = np.arange(N + 1) * h
    Zc, Yc, Xc = np.meshgrid(xs, xs, xs, indexing="ij")
    phi = levelset(Xc, Yc, Zc)  # (N+1)^3 indexed [k, j, i]
    corner = np.stack([phi[(c >> 2) & 1 : N + ((c >> 2) & 1), (c >> 1) & 1 : N + ((c >> 1) & 1), c & 1 : N + (c & 1)]
                       for c in range(8)], axis=-1)  # (N,N,N,8)
    n_in = (corner >= 0.0).sum(axis=-1)
    full = n_in == 8
    cut = (n_in > 0) & (n_in < 8)

    H = Stencil3D(nb, nb, nb, p)
    K = Stencil3D(nb, nb, nb, p)
    # Full cells: for every local pair combination the contribution over all
    # cells is a masked separable 3D array, added to the offset's slab at the
    # dof shift (az, ay, ax): pure slicing, no scatter.
    mask = full.astype(float)
    Hv = H.data.reshape(-1, nb, nb, nb)
    Kv = K.data.reshape(-1, nb, nb, nb)
    if full.any():
        for az in range(p1):
            for bz in range(p1):
                mz = Ms[:, az, bz][:, None, None]
                kz = Ks[:, az, bz][:, None, None]
                for ay in range(p1):
                    for by in range(p1):
                        my = Ms[:, ay, by][None, :, None]
                        ky = Ks[:, ay, by][None, :, None]
                        mzy = mask * (mz * my)
                        kzy = mask * (kz * my + mz * ky)
                        for ax in range(p1):
                            for bx in range(p1):
                                o = H.offset_index(bz - az, by - ay, bx - ax)
                                mx = Ms[:, ax, bx][None, None, :]
                                kx = Ks[:, ax, bx][None, None, :]
                                Hv[o, az : az + N, ay : ay + N, ax : ax + N] += mzy * mx
                                Kv[o, az : az + N, ay : ay + N, ax : ax + N] += kzy * mx + mzy * kx

    # Cut cells: Kuhn tetrahedra clipped to the linear interpolant of phi.
    cz, cy, cx = np.nonzero(cut)
    cell_fraction = full.astype(float)
    ref_pts, ref_w = _tet_rule(q)
    nq = ref_w.size
    loc = np.arange(p1)
    LZ, LY, LX = np.meshgrid(loc, loc, loc, indexing="ij")
    LZ, LY, LX = LZ.ravel(), LY.ravel(), LX.ravel()
    OFF = H.offset_index(LZ[None, :] - LZ[:, None], LY[None, :] - LY[:, None], LX[None, :] - LX[:, None])  # (64,64)
    for s in range(0, cz.size, chunk_cells):
        ez, ey, ex = cz[s : s + chunk_cells], cy[s : s + chunk_cells], cx[s : s + chunk_cells]
        nc = ez.size
        base = np.stack([ex, ey, ez], axis=1) * h  # (nc, 3) as (x, y, z)
        P = base[:, None, None, :] + h * _CORNERS[_KUHN][None, :, :, :]  # (nc, 6, 4, 3)
        Fc = corner[ez, ey, ex]  # (nc, 8)
        F = Fc[:, _KUHN]  # (nc, 6, 4)
        P = P.reshape(nc * 6, 4, 3)
        F = F.reshape(nc * 6, 4)
        sub, parent = _clip_tets(P, F)
        cell_of = parent // 6
        e1 = sub[:, 1] - sub[:, 0]
        e2 = sub[:, 2] - sub[:, 0]
        e3 = sub[:, 3] - sub[:, 0]
        det = np.abs(np.einsum("ti,ti->t", e1, np.cross(e2, e3)))
        vol = det / 6.0
        frac = np.bincount(cell_of, weights=vol, minlength=nc) / h**3
        cell_fraction[ez, ey, ex] = frac
        pts = (sub[:, 0, None, :] + ref_pts[None, :, 0, None] * e1[:, None, :] + ref_pts[None, :, 1, None] * e2[:, None, :]
               + ref_pts[None, :, 2, None] * e3[:, None, :])  # (S, nq, 3)
        wts = det[:, None] * ref_w[None, :]
        # Group sub-tets by cell: per cell, points = concatenation of its sub-tets.
        order = np.argsort(cell_of, kind="stable")
        cell_sorted = cell_of[order]
        counts = np.bincount(cell_sorted, minlength=nc)
        maxc = int(counts.max()) if counts.size else 0
        if maxc == 0:
            continue
        slot = np.arange(order.size) - np.repeat(np.cumsum(counts) - counts, counts)
        PP = np.zeros((nc, maxc, nq, 3))
        WW = np.zeros((nc, maxc, nq))
        PP[cell_sorted, slot] = pts[order]
        WW[cell_sorted, slot] = wts[order]
        # Points of empty slots sit at the cell's origin with zero weight.
        PP += (np.arange(maxc)[None, :, None, None] >= counts[:, None, None, None]) * base[:, None, None, :]
        PP = PP.reshape(nc, maxc * nq, 3)
        WW = WW.reshape(nc, maxc * nq)
        npt = maxc * nq
        elem = [np.repeat(ex, npt), np.repeat(ey, npt), np.repeat(ez, npt)]
        vals = []
        ders = []
        for ax in range(3):
            xq = np.clip(PP[:, :, ax].ravel(), elem[ax] * h, (elem[ax] + 1) * h)
            V, D = local_values(bases[ax], elem[ax], xq)
            vals.append(V.reshape(p1, nc, npt))
            ders.append(D.reshape(p1, nc, npt))
        Vx, Vy, Vz = vals
        Dx, Dy, Dz = ders
        def tensor(a, b, c):  # (p1,nc,npt) x3 -> (nc, p1^3, npt), row (z, y, x)
            t = a[:, None, None] * b[None, :, None] * c[None, None, :]
            return t.reshape(p1**3, nc, npt).transpose(1, 0, 2)

        W = WW[:, None, :]
        phi_ = tensor(Vz, Vy, Vx)
        Mloc = np.matmul(phi_ * W, phi_.transpose(0, 2, 1))
        Kloc = np.zeros_like(Mloc)
        for g in (tensor(Vz, Vy, Dx), tensor(Vz, Dy, Vx), tensor(Dz, Vy, Vx)):
            Kloc += np.matmul(g * W, g.transpose(0, 2, 1))
        rows = (((ez[:, None] + LZ[None, :]) * nb + (ey[:, None] + LY[None, :])) * nb + (ex[:, None] + LX[None, :]))  # (nc,64)
        R = np.repeat(rows[:, :, None], p1**3, axis=2)
        O = np.broadcast_to(OFF[None], R.shape)
        np.add.at(H.data, (O.ravel(), R.ravel()), Mloc.ravel())
        np.add.at(K.data, (O.ravel(), R.ravel()), Kloc.ravel())

    Hc = H.tocsr()
    Kc = K.tocsr()
    Hc = (0.5 * (Hc + Hc.T)).tocsr()
    Kc = (0.5 * (Kc + Kc.T)).tocsr()
    m1 = np.zeros(nb)
    for e in range(N):
        m1[e : e + p1] += np.diag(Ms[e])
    hbg = np.einsum("k,j,i->kji", m1, m1, m1).reshape(-1)
    diag = Hc.diagonal()
    gamma = np.zeros(nb**3)
    pos = hbg > 0
    gamma[pos] = diag[pos] / hbg[pos]
    active = np.nonzero(gamma > 0.0)[0]
    return CutAssembly3D(bases=bases, H=Hc, K=Kc, h_background_diag=hbg, gamma=gamma, active=active,
                         cell_fraction=cell_fraction, n_cut_cells=int(cz.size))


# ---------------------------------------------------------------------------
# Aggregation and restriction.
# ---------------------------------------------------------------------------


def greville_3d(bases) -> np.ndarray:
    gx, gy, gz = (b.greville for b in bases)
    nx, ny, nz = (b.n_basis for b in bases)
    X = np.tile(np.tile(gx, ny), nz)
    Y = np.tile(np.repeat(gy, nx), nz)
    Z = np.repeat(gz, nx * ny)
    return np.column_stack([X, Y, Z])


def monomial_rows_3d(bases, degree: int) -> np.ndarray:
    cx = polynomial_coefficients(bases[0], degree)
    cy = polynomial_coefficients(bases[1], degree)
    cz = polynomial_coefficients(bases[2], degree)
    rows = []
    for total in range(degree + 1):
        for a in range(total + 1):
            for b in range(total - a + 1):
                c = total - a - b
                rows.append(np.einsum("k,j,i->kji", cz[c], cy[b], cx[a]).reshape(-1))
    return np.vstack(rows)


@dataclass
class Aggregation3D:
    E: sp.csr_matrix
    interior: np.ndarray
    small_support: np.ndarray
    active: np.ndarray
    gamma: np.ndarray
    polynomial_residual: float


def _nearest_ordered(greville, interior_mask, lines, b, need, shape):
    """Interior dofs ordered by Greville distance from ``b`` (ties by index).

    Searches a growing index-space window; the window is accepted once its
    ``need``-th nearest point is closer than any point outside the window.
    """
    nz, ny, nx = shape
    kz, ky, kx = np.unravel_index(b, shape)
    gb = greville[b]
    r = 3
    while True:
        z0, z1 = max(kz - r, 0), min(kz + r + 1, nz)
        y0, y1 = max(ky - r, 0), min(ky + r + 1, ny)
        x0, x1 = max(kx - r, 0), min(kx + r + 1, nx)
        Z, Y, X = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
        idx = ((Z * ny + Y) * nx + X).ravel()
        idx = idx[interior_mask[idx]]
        idx.sort()
        whole = z0 == 0 and y0 == 0 and x0 == 0 and z1 == nz and y1 == ny and x1 == nx
        if idx.size >= need or whole:
            dist = np.linalg.norm(greville[idx] - gb, axis=1)
            order = np.argsort(dist, kind="stable")
            if whole:
                return idx[order]
            # Smallest distance a point outside the window could have.
            bound = np.inf
            for axis, (lo, hi, k) in enumerate(((x0, x1, kx), (y0, y1, ky), (z0, z1, kz))):
                line = lines[axis]
                if lo > 0:
                    bound = min(bound, gb[axis] - line[lo - 1])
                if hi < line.size:
                    bound = min(bound, line[hi] - gb[axis])
            if idx.size >= need and dist[order[need - 1]] < bound:
                return idx[order]
        r *= 2


def extension_core(gamma, greville, poly, shape, lines, *, support_threshold=0.04, tie_order="stable"):
    """The 3D extension rule on given Greville points and monomial rows.

    Returns ``(E, interior, small, active, polynomial_residual)``.
    """
    active = np.nonzero(gamma > 0.0)[0]
    interior = np.nonzero(gamma >= support_threshold)[0]
    small = active[gamma[active] < support_threshold]
    if interior.size == 0:
        raise ValueError("no interior coordinates")
    n_full = gamma.size
    n_terms = poly.shape[0]
    col = np.full(n_full, -1, dtype=np.int64)
    col[interior] = np.arange(interior.size)
    imask = np.zeros(n_full, dtype=bool)
    imask[interior] = True
    rows_t = [interior.astype(np.int64)]
    cols_t = [np.arange(interior.size, dtype=np.int64)]
    vals_t = [np.ones(interior.size)]
    cap = min(interior.size, 6 * n_terms)
    g_int = greville[interior]
    for b in small:
        if tie_order == "numpy-default":
            ordered = interior[np.argsort(np.linalg.norm(g_int - greville[int(b)], axis=1))]
        else:
            ordered = _nearest_ordered(greville, imask, lines, int(b), min(cap + 8, interior.size), shape)
        k = n_terms
        chosen = ordered[: k + 8]
        C = poly[:, chosen]
        while np.linalg.matrix_rank(C, tol=1.0e-10) < n_terms and len(chosen) < min(len(ordered), 6 * n_terms):
            k += 8
            chosen = ordered[: k + 8]
            C = poly[:, chosen]
        w, *_ = np.linalg.lstsq(C, poly[:, int(b)], rcond=None)
        keep = np.abs(w) > 1.0e-14
        rows_t.append(np.full(int(keep.sum()), int(b), dtype=np.int64))
        cols_t.append(col[chosen[keep]])
        vals_t.append(w[keep])
    E = sp.csr_matrix((np.concatenate(vals_t), (np.concatenate(rows_t), np.concatenate(cols_t))),
                      shape=(n_full, interior.size))
    active_mask = (gamma > 0.0).astype(float)
    res = 0.0
    for row in poly:
        res = max(res, float(np.max(np.abs(E @ row[interior] - row * active_mask))))
    return E, interior, small, active, res


def polynomial_preserving_extension_3d(asm: CutAssembly3D, *, support_threshold: float = 0.04,
                                       preserve_degree: int | None = None, tie_order: str = "stable") -> Aggregation3D:
    """Sparse ``E`` (``n_full x n_interior``); see the module docstring for the rule.

    ``tie_order="stable"`` breaks equal Greville distances by dof index
    (windowed nearest-neighbour search); ``"numpy-default"`` sorts the full
    interior distance row with NumPy's default (unstable) ``argsort`` as the
    reference does, whose tie order is that of this host's NumPy build.
    """
    bases = asm.bases
    degree = bases[0].degree if preserve_degree is None else int(preserve_degree)
    shape = (bases[2].n_basis, bases[1].n_basis, bases[0].n_basis)
    poly = monomial_rows_3d(bases, degree)
    greville = greville_3d(bases)
    lines = [b_.greville for b_ in bases]
    E, interior, small, active, res = extension_core(asm.gamma, greville, poly, shape, lines,
                                                     support_threshold=support_threshold, tie_order=tie_order)
    return Aggregation3D(E=E, interior=interior, small_support=small, active=active, gamma=asm.gamma,
                         polynomial_residual=res)


def restrict_spd(A: sp.csr_matrix, E: sp.csr_matrix) -> sp.csr_matrix:
    """``sym(E^T A E)`` (``code/bell_3d_core.py:550-552``)."""
    B = (E.T @ A @ E).tocsr()
    return (0.5 * (B + B.T)).tocsr()


@dataclass
class System3D:
    A: sp.csr_matrix
    E: sp.csr_matrix
    interior: np.ndarray
    small_support: np.ndarray
    active: np.ndarray
    meta: dict


def build_system_3d(levelset, n_elements: int, degree: int, *, quadrature_order: int | None = None,
                    support_threshold: float = 0.04) -> tuple[System3D, CutAssembly3D, Aggregation3D]:
    asm = assemble_levelset_3d(levelset, n_elements, degree, quadrature_order)
    agg = polynomial_preserving_extension_3d(asm, support_threshold=support_threshold)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    meta = dict(n_elements=n_elements, degree=degree, n_full=asm.n_full, n_active=int(asm.active.size),
                n_reduced=int(A.shape[0]), nnz=int(A.nnz), n_cut_cells=asm.n_cut_cells,
                min_cut_fraction=asm.min_cut_fraction, polynomial_residual=agg.polynomial_residual)
    return System3D(A, agg.E, agg.interior, agg.small_support, agg.active, meta), asm, agg


def planar_grid_for_target(cut: PlanarCut, degree: int, target: int, lo: int = 8, hi: int = 160) -> int:
    """Smallest-error element count whose fluid-side Greville count is ``target``."""
    best, best_err = lo, None
    for N in range(lo, hi + 1):
        g = make_open_uniform_basis(N, degree).greville
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        cnt = int((cut(X, Y, Z) >= 0).sum())
        err = abs(cnt - target)
        if best_err is None or err < best_err:
            best, best_err = N, err
        if cnt > 2 * target:
            break
    return best


def planar_cut_system(seed: int = 0, degree: int = 3, target_unknowns: int = 81_000, n_elements: int | None = None,
                      quadrature_order: int | None = None):
    """3D cut-box screened Poisson with a seeded random planar cut."""
    cut = PlanarCut.random(seed)
    N = n_elements or planar_grid_for_target(cut, degree, target_unknowns)
    system, asm, agg = build_system_3d(cut, N, degree, quadrature_order=quadrature_order)
    system.meta.update(kind="planar3d", seed=seed, normal=cut.normal.tolist(), offset=cut.offset)
    return system
