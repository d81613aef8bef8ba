"""Anisotropic 3D tensor B-spline grids (clamped or periodic axes) and their
cut-cell data for the matrix-free operator.

An ``Axis`` is a uniform grid of ``N`` elements of size ``L/N``. Clamped axes
carry the open-uniform basis of ``cutcell.basis`` (``N + p`` functions, the
reference's ``code/cut_spline_sbp.py:52-76`` scaled to ``[0, L]``); periodic
axes carry ``N`` uniform B-splines wrapped around the axis. Either way, on
element ``e`` the ``p + 1`` live functions are ``e + a`` (``a = 0..p``; modulo
``N`` on a periodic axis), which is the index convention the HIP kernels use.

``levelset_cut_data`` produces, for a fluid region ``phi >= 0``, the per-cell
classification, the Gram blocks of the cut cells (Kuhn-tetrahedra clipping of
the linear interpolant of ``phi``, as in ``cutcell.cut3d``), ``gamma`` and the
active / interior / small sets, without assembling any global matrix.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from .basis import gauss_legendre, make_open_uniform_basis
from .cut3d import _CORNERS, _KUHN, _clip_tets, _local_basis_degree, _tet_rule


@dataclass
class Axis:
    n_elements: int
    degree: int
    length: float = 1.0
    periodic: bool = False
    knots: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        N, p = self.n_elements, self.degree
        if self.periodic:
            if N < 2 * p + 1:
                raise ValueError("a periodic axis needs at least 2p+1 elements")
            self.knots = (np.arange(N + 2 * p + 1) - p) / N  # parametric, extended
        else:
            self.knots = make_open_uniform_basis(N, p).knots

    @property
    def n_basis(self) -> int:
        return self.n_elements if self.periodic else self.n_elements + self.degree

    @property
    def h(self) -> float:
        return self.length / self.n_elements

    def local_values(self, elements: np.ndarray, x: np.ndarray):
        """Values and d/dx of the local functions ``e + a`` at physical ``x``."""
        p = self.degree
        e = np.asarray(elements, dtype=np.int64)
        t = np.asarray(x, dtype=float) / self.length
        span = e + p
        V = _local_basis_degree(self.knots, span, t, p)
        L = _local_basis_degree(self.knots, span, t, p - 1)
        D = np.zeros_like(V)
        kn = self.knots
        for a in range(p + 1):
            i = span - p + a
            if a >= 1:
                den = kn[i + p] - kn[i]
                D[a] += np.where(den > 0, p * L[a - 1] / np.where(den > 0, den, 1.0), 0.0)
            if a <= p - 1:
                den = kn[i + p + 1] - kn[i + 1]
                D[a] -= np.where(den > 0, p * L[a] / np.where(den > 0, den, 1.0), 0.0)
        return V, D / self.length

    def element_forms(self, order: int | None = None):
        """Per-element mass and stiffness, ``(N, p+1, p+1)`` (physical scaling)."""
        p = self.degree
        q = order or max(p + 2, 5)
        g, w = gauss_legendre(q)
        N = self.n_elements
        h = self.h
        e = np.repeat(np.arange(N), q)
        x = ((np.arange(N)[:, None] + 0.5 + 0.5 * g[None, :]) * h).ravel()
        wq = np.tile(0.5 * h * w, N)
        V, D = self.local_values(e, x)
        V = V.reshape(p + 1, N, q)
        D = D.reshape(p + 1, N, q)
        W = wq.reshape(N, q)
        return np.einsum("aeq,eq,beq->eab", V, W, V), np.einsum("aeq,eq,beq->eab", D, W, D)

    def dof(self, e, a):
        j = np.asarray(e) + np.asarray(a)
        return j % self.n_elements if self.periodic else j

    def band(self, forms):
        """Global banded matrix from element forms: ``B[i, d]`` = entry ``(i, i+d-p)``."""
        p = self.degree
        nb = self.n_basis
        B = np.zeros((nb, 2 * p + 1))
        for e in range(self.n_elements):
            for a in range(p + 1):
                for b in range(p + 1):
                    B[self.dof(e, a), b - a + p] += forms[e, a, b]
        return B

    def greville(self) -> np.ndarray:
        p = self.degree
        if self.periodic:
            # Function j lives on [(j-p)h, (j+1)h]; its Greville point is the
            # mean of its p interior knots, kept unwrapped in [0, L) + shift.
            j = np.arange(self.n_basis)
            return ((j - p + 1 + (p - 1) / 2.0) * self.h) % self.length
        b = make_open_uniform_basis(self.n_elements, p)
        return b.greville * self.length

    def monomial_coefficients(self, max_degree: int) -> np.ndarray:
        """Spline coefficients of ``x^r`` (Marsden's identity on the knots).

        For a periodic axis the identity holds on any window that does not
        wrap; coefficients are expressed in the unwrapped coordinate.
        """
        p = self.degree
        nb = self.n_basis
        kn = self.knots * self.length
        out = np.zeros((max_degree + 1, nb))
        from itertools import combinations
        from math import comb

        for j in range(nb):
            inner = kn[j + 1 : j + p + 1]  # the p interior knots of function j
            for r in range(max_degree + 1):
                if r > p:
                    continue
                s = sum(np.prod(c) for c in combinations(inner, r)) if r > 0 else 1.0
                out[r, j] = s / comb(p, r)
        return out


@dataclass
class Grid3D:
    axes: tuple  # (x, y, z)

    @property
    def degree(self) -> int:
        return self.axes[0].degree

    @property
    def nb(self):
        return tuple(a.n_basis for a in self.axes)

    @property
    def ne(self):
        return tuple(a.n_elements for a in self.axes)

    @property
    def n_full(self) -> int:
        return int(np.prod(self.nb))

    def dof_index(self, ix, iy, iz):
        nbx, nby, _ = self.nb
        return (np.asarray(iz) * nby + np.asarray(iy)) * nbx + np.asarray(ix)


@dataclass
class CutData:
    grid: Grid3D
    cell_class: np.ndarray  # (nez, ney, nex): 0 empty, 1 full, 2 cut
    cut_cells: np.ndarray  # (n_cut, 3) as (ex, ey, ez)
    cut_mass: np.ndarray  # (n_cut, (p+1)^3, (p+1)^3)
    cut_stiff: np.ndarray
    cut_fraction: np.ndarray  # (n_cut,)
    gamma: np.ndarray  # (n_full,)


def _gram_chunk(args):
    """Gram blocks of a chunk of cut cells from their clipped sub-tetrahedra."""
    from threadpoolctl import threadpool_limits

    with threadpool_limits(limits=1):  # one BLAS thread per pool worker
        return _gram_chunk_body(*args)


def _gram_chunk_body(axes_spec, p, q, cells, sub, cell_of):
    axes = [Axis(*spec) for spec in axes_spec]
    p1 = p + 1
    nc = cells.shape[0]
    ref_pts, ref_w = _tet_rule(q)
    nq = ref_w.size
    hvec = np.array([a.h for a in axes])
    base = cells * hvec
    e1 = sub[:, 1] - sub[:, 0]
    e2 = sub[:, 2] - sub[:, 0]
    e3 = sub[:, 3] - sub[:, 0]
    det = np.abs(np.einsum("ti,ti->t", e1, np.cross(e2, e3)))
    frac = np.bincount(cell_of, weights=det / 6.0, minlength=nc) / np.prod(hvec)
    pts = (sub[:, 0, None, :] + ref_pts[None, :, 0, None] * e1[:, None, :] + ref_pts[None, :, 1, None] * e2[:, None, :]
           + ref_pts[None, :, 2, None] * e3[:, None, :])
    wts = det[:, None] * ref_w[None, :]
    order = np.argsort(cell_of, kind="stable")
    cs = cell_of[order]
    counts = np.bincount(cs, minlength=nc)
    maxc = int(counts.max())
    slot = np.arange(order.size) - np.repeat(np.cumsum(counts) - counts, counts)
    PP = np.zeros((nc, maxc, nq, 3))
    WW = np.zeros((nc, maxc, nq))
    PP[cs, slot] = pts[order]
    WW[cs, slot] = wts[order]
    PP += (np.arange(maxc)[None, :, None, None] >= counts[:, None, None, None]) * base[:, None, None, :]
    npt = maxc * nq
    PP = PP.reshape(nc, npt, 3)
    WW = WW.reshape(nc, npt)
    vals, ders = [], []
    for k in range(3):
        axk = axes[k]
        e_rep = np.repeat(cells[:, k], npt)
        xq = np.clip(PP[:, :, k].ravel(), e_rep * axk.h, (e_rep + 1) * axk.h)
        V, D = axk.local_values(e_rep, xq)
        vals.append(V.reshape(p1, nc, npt))
        ders.append(D.reshape(p1, nc, npt))
    Vx, Vy, Vz = vals
    Dx, Dy, Dz = ders

    def tensor(a, b, c):
        return (a[:, None, None] * b[None, :, None] * c[None, None, :]).reshape(p1**3, nc, npt).transpose(1, 0, 2)

    W = WW[:, None, :]
    phi_ = tensor(Vz, Vy, Vx)
    mass = np.matmul(phi_ * W, phi_.transpose(0, 2, 1))
    stiff = np.zeros_like(mass)
    for g in (tensor(Vz, Vy, Dx), tensor(Vz, Dy, Vx), tensor(Dz, Vy, Vx)):
        stiff += np.matmul(g * W, g.transpose(0, 2, 1))
    return mass, stiff, frac


class _Cut3dDesc(ctypes.Structure):
    """``cc_cut3d_desc`` of the C ABI."""

    _fields_ = [
        ("n_cut", ctypes.c_int32),
        ("degree", ctypes.c_int32),
        ("ne", ctypes.c_int32 * 3),
        ("n_knots", ctypes.c_int32 * 3),
        ("length", ctypes.c_double * 3),
        ("knots", ctypes.c_void_p * 3),
        ("cells", ctypes.c_void_p),
        ("corner", ctypes.c_void_p),
        ("nq", ctypes.c_int32),
        ("rule_pts", ctypes.c_void_p),
        ("rule_w", ctypes.c_void_p),
    ]


def gram_blocks_device(grid: Grid3D, cells: np.ndarray, corner: np.ndarray, q: int):
    """The cut cells' Gram blocks on the GPU (``cc_cut3d_gram``): ``cells``
    (n, 3) as (ex, ey, ez), ``corner`` (n, 8) level-set values at corner
    c = x + 2 y + 4 z. Returns host ``(mass, stiff, frac)`` as the host
    integration does, in the order of ``cells``."""
    import torch

    from . import hip

    lib = hip.require_gpu()
    p1 = grid.degree + 1
    P = p1**3
    n = cells.shape[0]
    ref_pts, ref_w = _tet_rule(q)
    keep = []

    def dev(a, dt):
        t = torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to("cuda")
        keep.append(t)
        return t.data_ptr()

    d = _Cut3dDesc()
    d.n_cut = n
    d.degree = grid.degree
    for k, a in enumerate(grid.axes):
        d.ne[k] = a.n_elements
        d.n_knots[k] = a.knots.size
        d.length[k] = a.length
        d.knots[k] = dev(a.knots, np.float64)
    d.cells = dev(cells, np.int32)
    d.corner = dev(corner, np.float64)
    d.nq = ref_w.size
    d.rule_pts = dev(ref_pts, np.float64)
    d.rule_w = dev(ref_w, np.float64)
    mass = torch.empty((n, P, P), dtype=torch.float64, device="cuda")
    stiff = torch.empty_like(mass)
    frac = torch.empty(n, dtype=torch.float64, device="cuda")
    hip._check(lib.cc_cut3d_gram(ctypes.addressof(d), hip._ptr(mass), hip._ptr(stiff), hip._ptr(frac), hip._stream()),
               "cut3d_gram")
    out = mass.cpu().numpy(), stiff.cpu().numpy(), frac.cpu().numpy()
    del keep
    return out


def levelset_cut_data(grid: Grid3D, levelset, quadrature_order: int | None = None, chunk_cells: int = 32,
                      jobs: int | None = None, device: bool = False) -> CutData:
    """Classify cells and integrate the cut cells' local Gram blocks.

    Cut cells are clipped all at once, sorted by their number of
    sub-tetrahedra (so chunks carry little padding) and integrated in a
    process pool of ``jobs`` workers (default: the CPU count, at most 16);
    with ``device`` the integration runs on the GPU instead, one workgroup
    per cut cell (``gram_blocks_device``).
    """
    import os
    from concurrent.futures import ProcessPoolExecutor

    ax, ay, az = grid.axes
    p = grid.degree
    p1 = p + 1
    q = quadrature_order or (p + 3)
    nex, ney, nez = grid.ne
    xs = np.arange(nex + 1) * ax.h
    ys = np.arange(ney + 1) * ay.h
    zs = np.arange(nez + 1) * az.h
    Zc, Yc, Xc = np.meshgrid(zs, ys, xs, indexing="ij")
    phi = levelset(Xc, Yc, Zc)
    corner = np.stack(
        [phi[(c >> 2) & 1 : nez + ((c >> 2) & 1), (c >> 1) & 1 : ney + ((c >> 1) & 1), c & 1 : nex + (c & 1)] for c in range(8)],
        axis=-1,
    )
    n_in = (corner >= 0.0).sum(axis=-1)
    cls = np.where(n_in == 8, 1, np.where(n_in == 0, 0, 2)).astype(np.int8)
    cz, cy, cx = np.nonzero(cls == 2)
    hvec = np.array([ax.h, ay.h, az.h])
    nc_all = cz.size
    if nc_all and device:
        cells = np.stack([cx, cy, cz], axis=1).astype(np.int64)
        m, st, fr = gram_blocks_device(grid, cells, corner[cz, cy, cx], q)
        masses, stiffs, fracs = [m], [st], [fr]
    elif nc_all:
        base = np.stack([cx, cy, cz], axis=1) * hvec
        P = base[:, None, None, :] + hvec * _CORNERS[_KUHN][None, :, :, :]
        F = corner[cz, cy, cx][:, _KUHN]
        sub_all, parent = _clip_tets(P.reshape(nc_all * 6, 4, 3), F.reshape(nc_all * 6, 4))
        cell_all = parent // 6
        nsub = np.bincount(cell_all, minlength=nc_all)
        by_count = np.argsort(nsub, kind="stable")
        rank = np.empty(nc_all, dtype=np.int64)
        rank[by_count] = np.arange(nc_all)
        order = np.argsort(rank[cell_all], kind="stable")
        sub_sorted = sub_all[order]
        owner = rank[cell_all[order]]  # cell position in count order
        starts = np.searchsorted(owner, np.arange(0, nc_all, chunk_cells))
        ends = np.append(starts[1:], owner.size)
        axes_spec = [(a.n_elements, a.degree, a.length, a.periodic) for a in grid.axes]
        cells_sorted = np.stack([cx, cy, cz], axis=1)[by_count].astype(np.int64)
        tasks = []
        for k, c0 in enumerate(range(0, nc_all, chunk_cells)):
            c1 = min(c0 + chunk_cells, nc_all)
            tasks.append((axes_spec, p, q, cells_sorted[c0:c1], sub_sorted[starts[k] : ends[k]],
                          owner[starts[k] : ends[k]] - c0))
        jobs = jobs or min(16, os.cpu_count() or 1)
        if jobs > 1 and len(tasks) > 1:
            with ProcessPoolExecutor(max_workers=jobs) as ex:
                parts = list(ex.map(_gram_chunk, tasks))
        else:
            parts = [_gram_chunk(t) for t in tasks]
        masses = [m for m, _, _ in parts]
        stiffs = [s_ for _, s_, _ in parts]
        fracs = [f for _, _, f in parts]
        # Back to the row-major order of the cut cells.
        inv = np.empty(nc_all, dtype=np.int64)
        inv[by_count] = np.arange(nc_all)
        mass_sorted = np.concatenate(masses)
        stiff_sorted = np.concatenate(stiffs)
        frac_sorted = np.concatenate(fracs)
        masses, stiffs, fracs = [mass_sorted[inv]], [stiff_sorted[inv]], [frac_sorted[inv]]
    else:
        masses, stiffs, fracs = [], [], []
    n_cut = cz.size
    cut_mass = np.concatenate(masses) if masses else np.zeros((0, p1**3, p1**3))
    cut_stiff = np.concatenate(stiffs) if stiffs else np.zeros((0, p1**3, p1**3))
    cut_fraction = np.concatenate(fracs) if fracs else np.zeros(0)
    cut_mass = 0.5 * (cut_mass + cut_mass.transpose(0, 2, 1))
    cut_stiff = 0.5 * (cut_stiff + cut_stiff.transpose(0, 2, 1))

    # gamma = diag(H_fluid) / diag(H_background), per dof.
    forms = [a.element_forms() for a in grid.axes]
    dM = [np.zeros(a.n_basis) for a in grid.axes]
    for k, a in enumerate(grid.axes):
        for e in range(a.n_elements):
            for l in range(p1):
                dM[k][a.dof(e, l)] += forms[k][0][e, l, l]
    hbg = np.einsum("k,j,i->kji", dM[2], dM[1], dM[0]).reshape(-1)
    hfl = np.zeros(grid.n_full)
    # Full cells: add the tensor diagonal of every full cell.
    fz, fy, fx = np.nonzero(cls == 1)
    loc = np.arange(p1)
    for lz in range(p1):
        for ly in range(p1):
            for lx in range(p1):
                val = forms[2][0][fz, lz, lz] * forms[1][0][fy, ly, ly] * forms[0][0][fx, lx, lx]
                g = grid.dof_index(ax.dof(fx, lx), ay.dof(fy, ly), az.dof(fz, lz))
                np.add.at(hfl, g, val)
    if n_cut:
        LZ, LY, LX = np.meshgrid(loc, loc, loc, indexing="ij")
        gl = grid.dof_index(ax.dof(cx[:, None], LX.ravel()[None]), ay.dof(cy[:, None], LY.ravel()[None]),
                            az.dof(cz[:, None], LZ.ravel()[None]))
        np.add.at(hfl, gl.ravel(), np.einsum("cll->cl", cut_mass).ravel())
    gamma = np.where(hbg > 0, hfl / np.where(hbg > 0, hbg, 1.0), 0.0)
    return CutData(grid=grid, cell_class=cls, cut_cells=np.stack([cx, cy, cz], axis=1).astype(np.int64),
                   cut_mass=cut_mass, cut_stiff=cut_stiff, cut_fraction=cut_fraction, gamma=gamma)
