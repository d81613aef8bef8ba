"""Host preparation of the matrix-free aggregated operator ``E^T (K + c M) E``.

The operator is never assembled. On the device (``csrc/sumfact.hip``):

1. ``u = E x`` on the background dof grid (interior dofs copy ``x``, small
   dofs are their stencil combinations, inactive dofs are zero);
2. ``v = A_all u`` with the tensor-product operator of *all* background cells,
   ``A_all = Kz(x)My(x)Mx + Mz(x)Ky(x)Mx + Mz(x)My(x)Kx + c Mz(x)My(x)Mx``,
   applied by sum factorisation: three passes of banded 1D operators staged in
   LDS;
3. corrections on the thin shell of cells that are not fully fluid but touch
   an active dof: ``v += (C_e - A_e) u_e`` for a cut cell (``C_e`` its
   quadrature Gram block) and ``v -= A_e u_e`` for an empty one; cells whose
   dofs are all inactive only ever see zeros and feed discarded outputs;
4. ``y = E^T v`` (interior outputs plus the transposed small-dof stencils).

This module computes the shell, the cut blocks, ``E`` (the reference's 3D
extension rule, ``code/bell_3d_core.py:471-542``) and the exact Jacobi
diagonal of ``E^T A E``, and packs them for the C ABI (``cc_sumfact_desc``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .cut3d import extension_core
from .grid3d import CutData, Grid3D, levelset_cut_data


@dataclass
class MatrixFreeSystem:
    grid: Grid3D
    mass_coef: float
    cut: CutData
    E: sp.csr_matrix
    interior: np.ndarray
    small: np.ndarray
    active: np.ndarray
    corr_cells: np.ndarray  # (n_corr, 3) (ex, ey, ez)
    corr_block: np.ndarray  # (n_corr,) index into blocks, -1 for an empty cell
    blocks: np.ndarray  # (n_cut, P, P): C_e = K_cut + c M_cut
    touch_dof: np.ndarray
    touch_ptr: np.ndarray
    touch_idx: np.ndarray
    diag: np.ndarray
    elem_m: list
    elem_k: list
    band_m: list
    band_k: list
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.interior.size)


def greville_grid(grid: Grid3D) -> np.ndarray:
    gx, gy, gz = (a.greville() for a in grid.axes)
    nx, ny, nz = grid.nb
    return np.column_stack([np.tile(np.tile(gx, ny), nz), np.tile(np.repeat(gy, nx), nz), np.repeat(gz, nx * ny)])


def monomial_rows(grid: Grid3D, degree: int) -> np.ndarray:
    cx, cy, cz = (a.monomial_coefficients(degree) for a in grid.axes)
    rows = []
    for total in range(degree + 1):
        for a in range(total + 1):
            for b in range(total - a + 1):
                c = total - a - b
                rows.append(np.einsum("k,j,i->kji", cz[c], cy[b], cx[a]).reshape(-1))
    return np.vstack(rows)


def _cell_dofs(grid: Grid3D, cells: np.ndarray) -> np.ndarray:
    """(n, P) full dof ids of the local functions of each cell, local order (z, y, x)."""
    p1 = grid.degree + 1
    ax, ay, az = grid.axes
    loc = np.arange(p1)
    LZ, LY, LX = np.meshgrid(loc, loc, loc, indexing="ij")
    return grid.dof_index(ax.dof(cells[:, 0:1], LX.ravel()[None]), ay.dof(cells[:, 1:2], LY.ravel()[None]),
                          az.dof(cells[:, 2:3], LZ.ravel()[None]))


def _full_cell_blocks(grid: Grid3D, cells: np.ndarray, elem_m, elem_k, c: float) -> np.ndarray:
    """Dense ``A_e`` of uncut cells, local order (z, y, x)."""
    Mx, My, Mz = (elem_m[k][cells[:, k]] for k in range(3))
    Kx, Ky, Kz = (elem_k[k][cells[:, k]] for k in range(3))
    kr = lambda A, B, C: np.einsum("nab,ncd,nef->nacebdf", A, B, C).reshape(cells.shape[0], -1, A.shape[1] ** 3)
    return kr(Kz, My, Mx) + kr(Mz, Ky, Mx) + kr(Mz, My, Kx) + c * kr(Mz, My, Mx)


def build_matrix_free(grid: Grid3D, levelset, *, mass_coef: float = 1.0, support_threshold: float = 0.04,
                      quadrature_order: int | None = None, device: bool = False) -> MatrixFreeSystem:
    """The matrix-free system of a level-set cut; ``device`` integrates the
    cut cells' Gram blocks on the GPU (``levelset_cut_data``)."""
    p = grid.degree
    p1 = p + 1
    cut = levelset_cut_data(grid, levelset, quadrature_order, device=device)
    shape = (grid.nb[2], grid.nb[1], grid.nb[0])
    poly = monomial_rows(grid, p)
    gre = greville_grid(grid)
    lines = [a.greville() for a in grid.axes]
    E, interior, small, active, res = extension_core(cut.gamma, gre, poly, shape, lines,
                                                     support_threshold=support_threshold)
    n_full = grid.n_full
    is_active = np.zeros(n_full, dtype=bool)
    is_active[active] = True

    # Correction shell: cut cells, and empty cells touching an active dof.
    ez, ey, ex = np.nonzero(cut.cell_class == 0)
    empty = np.stack([ex, ey, ez], axis=1).astype(np.int64)
    if empty.size:
        touches = is_active[_cell_dofs(grid, empty)].any(axis=1)
        empty = empty[touches]
    corr_cells = np.concatenate([cut.cut_cells, empty]) if empty.size else cut.cut_cells.copy()
    corr_block = np.concatenate([np.arange(cut.cut_cells.shape[0]), -np.ones(len(empty), dtype=np.int64)]).astype(np.int64)
    blocks = cut.cut_stiff + mass_coef * cut.cut_mass

    # Dofs receiving corrections: index (s * P + l) into the per-cell outputs.
    P = p1**3
    dofs = _cell_dofs(grid, corr_cells)  # (n_corr, P)
    flat_dof = dofs.ravel()
    flat_idx = np.arange(flat_dof.size, dtype=np.int64)
    order = np.lexsort((flat_idx, flat_dof))
    sd = flat_dof[order]
    touch_dof, start = np.unique(sd, return_index=True)
    touch_ptr = np.concatenate([start, [sd.size]]).astype(np.int64)
    touch_idx = flat_idx[order]

    forms = [a.element_forms() for a in grid.axes]
    elem_m = [f[0] for f in forms]
    elem_k = [f[1] for f in forms]
    band_m = [a.band(f[0]) for a, f in zip(grid.axes, forms)]
    band_k = [a.band(f[1]) for a, f in zip(grid.axes, forms)]

    mf = MatrixFreeSystem(grid=grid, mass_coef=mass_coef, cut=cut, E=E, interior=interior, small=small, active=active,
                          corr_cells=corr_cells, corr_block=corr_block, blocks=blocks, touch_dof=touch_dof,
                          touch_ptr=touch_ptr, touch_idx=touch_idx, diag=np.zeros(interior.size),
                          elem_m=elem_m, elem_k=elem_k, band_m=band_m, band_k=band_k)
    mf.diag = jacobi_diagonal(mf)
    mf.meta = dict(n_full=n_full, n_active=int(active.size), n_reduced=int(interior.size), n_small=int(small.size),
                   n_cut_cells=int(cut.cut_cells.shape[0]), n_corr_cells=int(corr_cells.shape[0]),
                   polynomial_residual=res,
                   min_cut_fraction=float(cut.cut_fraction.min()) if cut.cut_fraction.size else 1.0)
    return mf


# ---------------------------------------------------------------------------
# Entries of A = A_all + shell corrections (host, for the Jacobi diagonal and tests).
# ---------------------------------------------------------------------------


def _offset(axis, i, j):
    """Band offset d with A(i, j) = band[i, d]; None when outside the band."""
    p = axis.degree
    d = j - i
    if axis.periodic:
        n = axis.n_basis
        d = (d + p) % n - p
    if d < -p or d > p:
        return None
    return d + p


def a_all_entry(mf: MatrixFreeSystem, g: int, h: int) -> float:
    nbx, nby, _ = mf.grid.nb
    gz, gy, gx = g // (nbx * nby), (g // nbx) % nby, g % nbx
    hz, hy, hx = h // (nbx * nby), (h // nbx) % nby, h % nbx
    ax, ay, az = mf.grid.axes
    dx, dy, dz = _offset(ax, gx, hx), _offset(ay, gy, hy), _offset(az, gz, hz)
    if dx is None or dy is None or dz is None:
        return 0.0
    Mx, My, Mz = mf.band_m[0][gx, dx], mf.band_m[1][gy, dy], mf.band_m[2][gz, dz]
    Kx, Ky, Kz = mf.band_k[0][gx, dx], mf.band_k[1][gy, dy], mf.band_k[2][gz, dz]
    return Kz * My * Mx + Mz * Ky * Mx + Mz * My * Kx + mf.mass_coef * Mz * My * Mx


def shell_blocks(mf: MatrixFreeSystem) -> np.ndarray:
    """``D_e = C_e - A_e`` (cut) or ``-A_e`` (empty) for the shell cells."""
    Afull = _full_cell_blocks(mf.grid, mf.corr_cells, mf.elem_m, mf.elem_k, mf.mass_coef)
    D = -Afull
    cutm = mf.corr_block >= 0
    D[cutm] += mf.blocks[mf.corr_block[cutm]]
    return D


def cut_corrections(mf: MatrixFreeSystem) -> np.ndarray:
    """``C_e - A_e`` for each cut block, in block order: what a cut cell adds
    to the full-grid tensor apply (the device descriptor's ``blocks``)."""
    D = mf.blocks.copy()
    cutm = mf.corr_block >= 0
    D[mf.corr_block[cutm]] -= _full_cell_blocks(mf.grid, mf.corr_cells[cutm], mf.elem_m, mf.elem_k, mf.mass_coef)
    return D


def jacobi_diagonal(mf: MatrixFreeSystem) -> np.ndarray:
    """Exact ``diag(E^T A E)``.

    With ``A = sum_e A_e`` over cells (tensor block for a fluid cell, Gram block
    for a cut cell, zero for an empty cell), ``diag_j = sum_e (E_e^T A_e E_e)_jj``
    where ``E_e`` are the rows of ``E`` for the cell's dofs. For a column with
    no small-dof entries only the ``A_gg`` terms survive; the cells that hold a
    small dof add the rest of the quadratic form.
    """
    grid = mf.grid
    p = grid.degree
    D = shell_blocks(mf)
    dofs_shell = _cell_dofs(grid, mf.corr_cells)
    bz, by, bx = (mf.band_m[k][:, p] for k in (2, 1, 0))
    kz, ky, kx = (mf.band_k[k][:, p] for k in (2, 1, 0))
    adiag = (np.einsum("k,j,i->kji", kz, by, bx) + np.einsum("k,j,i->kji", bz, ky, bx)
             + np.einsum("k,j,i->kji", bz, by, kx) + mf.mass_coef * np.einsum("k,j,i->kji", bz, by, bx)).reshape(-1)
    np.add.at(adiag, dofs_shell.ravel(), np.einsum("nll->nl", D).ravel())
    diag = adiag[mf.interior].copy()
    if mf.small.size == 0:
        return diag
    n_full = grid.n_full
    col = -np.ones(n_full, dtype=np.int64)
    col[mf.interior] = np.arange(mf.interior.size)
    is_small = np.zeros(n_full, dtype=bool)
    is_small[mf.small] = True
    Erows = mf.E.tocsr()
    # Cells holding at least one small dof.
    nex, ney, nez = grid.ne
    Z, Y, X = np.meshgrid(np.arange(nez), np.arange(ney), np.arange(nex), indexing="ij")
    cells = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    cd = _cell_dofs(grid, cells)
    hold = is_small[cd].any(axis=1)
    cells, cd = cells[hold], cd[hold]
    # The cell blocks A_e: tensor for fluid cells, Gram for cut, zero for empty.
    cls = mf.cut.cell_class[cells[:, 2], cells[:, 1], cells[:, 0]]
    cut_index = {tuple(c): i for i, c in enumerate(mf.cut.cut_cells.tolist())}
    for s0 in range(0, cells.shape[0], 512):
        cc = cells[s0 : s0 + 512]
        blocks = _full_cell_blocks(grid, cc, mf.elem_m, mf.elem_k, mf.mass_coef)
        for i in range(cc.shape[0]):
            k = cls[s0 + i]
            if k == 0:
                continue
            Ae = blocks[i] if k == 1 else mf.blocks[cut_index[tuple(cc[i].tolist())]]
            d = cd[s0 + i]
            sub = Erows[d]
            if sub.nnz == 0:
                continue
            cols = np.unique(sub.indices)
            L = sub[:, cols].toarray()  # (P, m)
            q = ((Ae @ L) * L).sum(axis=0)
            # Remove what adiag already holds: A_e[a, a] for interior dofs a.
            for a in np.nonzero(col[d] >= 0)[0]:
                j = col[d[a]]
                q[np.searchsorted(cols, j)] -= Ae[a, a]
            diag[cols] += q
    return diag


def assemble_reduced_csr(mf: MatrixFreeSystem) -> sp.csr_matrix:
    """Explicit ``E^T A E`` (host, small grids: the reference for parity tests)."""
    grid = mf.grid
    p1 = grid.degree + 1
    n_full = grid.n_full
    rows, cols, vals = [], [], []
    # Separable part over all cells (every cell, full blocks).
    nex, ney, nez = grid.ne
    Z, Y, X = np.meshgrid(np.arange(nez), np.arange(ney), np.arange(nex), indexing="ij")
    cells = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    for s0 in range(0, cells.shape[0], 4096):
        c = cells[s0 : s0 + 4096]
        B = _full_cell_blocks(grid, c, mf.elem_m, mf.elem_k, mf.mass_coef)
        d = _cell_dofs(grid, c)
        rows.append(np.repeat(d, d.shape[1], axis=1).ravel())
        cols.append(np.tile(d, (1, d.shape[1])).ravel())
        vals.append(B.ravel())
    D = shell_blocks(mf)
    d = _cell_dofs(grid, mf.corr_cells)
    rows.append(np.repeat(d, d.shape[1], axis=1).ravel())
    cols.append(np.tile(d, (1, d.shape[1])).ravel())
    vals.append(D.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n_full, n_full))
    return (mf.E.T @ A @ mf.E).tocsr()


# ---------------------------------------------------------------------------
# Device packing for the C ABI.
# ---------------------------------------------------------------------------


class SumfactDesc(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int32 * 3),
        ("ne", ctypes.c_int32 * 3),
        ("degree", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("mass_coef", ctypes.c_double),
        ("band_m", ctypes.c_void_p * 3),
        ("band_k", ctypes.c_void_p * 3),
        ("elem_m", ctypes.c_void_p * 3),
        ("elem_k", ctypes.c_void_p * 3),
        ("n_corr", ctypes.c_int32),
        ("corr_cell", ctypes.c_void_p),
        ("corr_block", ctypes.c_void_p),
        ("blocks", ctypes.c_void_p),
        ("n_touch", ctypes.c_int32),
        ("touch_dof", ctypes.c_void_p),
        ("touch_ptr", ctypes.c_void_p),
        ("touch_idx", ctypes.c_void_p),
        ("n_red", ctypes.c_int32),
        ("n_small", ctypes.c_int32),
        ("red2full", ctypes.c_void_p),
        ("full2red", ctypes.c_void_p),
        ("full2small", ctypes.c_void_p),
        ("small_ptr", ctypes.c_void_p),
        ("small_col", ctypes.c_void_p),
        ("small_w", ctypes.c_void_p),
        ("small_dof", ctypes.c_void_p),
        ("et_ptr", ctypes.c_void_p),
        ("et_small", ctypes.c_void_p),
        ("et_w", ctypes.c_void_p),
        ("diag", ctypes.c_void_p),
        ("n_et_rows", ctypes.c_int32),
        ("et_rows", ctypes.c_void_p),
    ]


class DeviceMatrixFree:
    """Device arrays of a ``MatrixFreeSystem`` and its packed descriptor."""

    def __init__(self, mf: MatrixFreeSystem, device: str = "cuda"):
        import torch

        self.mf = mf
        keep = []

        def t(a, dt):
            x = torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
            keep.append(x)
            return x.data_ptr()

        g = mf.grid
        d = SumfactDesc()
        for k in range(3):
            d.nb[k] = g.nb[k]
            d.ne[k] = g.ne[k]
            d.periodic[k] = int(g.axes[k].periodic)
            d.band_m[k] = t(mf.band_m[k], np.float64)
            d.band_k[k] = t(mf.band_k[k], np.float64)
            d.elem_m[k] = t(mf.elem_m[k], np.float64)
            d.elem_k[k] = t(mf.elem_k[k], np.float64)
        d.degree = g.degree
        d.mass_coef = mf.mass_coef
        d.n_corr = mf.corr_cells.shape[0]
        d.corr_cell = t(mf.corr_cells, np.int32)
        d.corr_block = t(mf.corr_block, np.int32)
        d.blocks = t(cut_corrections(mf) if mf.blocks.size else np.zeros(1), np.float64)
        self._blocks64 = keep[-1]
        self._blocks32 = None
        d.n_touch = mf.touch_dof.size
        d.touch_dof = t(mf.touch_dof, np.int32)
        d.touch_ptr = t(mf.touch_ptr, np.int32)
        d.touch_idx = t(mf.touch_idx, np.int32)
        n_full = g.n_full
        full2red = -np.ones(n_full, dtype=np.int64)
        full2red[mf.interior] = np.arange(mf.interior.size)
        full2small = -np.ones(n_full, dtype=np.int64)
        full2small[mf.small] = np.arange(mf.small.size)
        Es = mf.E[mf.small].tocsr()
        EsT = Es.T.tocsr()  # (n_red, n_small)
        d.n_red = mf.interior.size
        d.n_small = mf.small.size
        d.red2full = t(mf.interior, np.int32)
        d.full2red = t(full2red, np.int32)
        d.full2small = t(full2small, np.int32)
        d.small_ptr = t(Es.indptr, np.int32)
        d.small_col = t(Es.indices if Es.nnz else np.zeros(1), np.int32)
        d.small_w = t(Es.data if Es.nnz else np.zeros(1), np.float64)
        d.small_dof = t(mf.small if mf.small.size else np.zeros(1), np.int32)
        d.et_ptr = t(EsT.indptr, np.int32)
        d.et_small = t(EsT.indices if EsT.nnz else np.zeros(1), np.int32)
        d.et_w = t(EsT.data if EsT.nnz else np.zeros(1), np.float64)
        d.diag = t(mf.diag, np.float64)
        et_rows = np.nonzero(np.diff(EsT.indptr))[0]
        d.n_et_rows = et_rows.size
        d.et_rows = t(et_rows if et_rows.size else np.zeros(1), np.int32)
        self.desc = d
        self._keep = keep
        self.n = mf.interior.size
        self.device = device

    # -- device operations ---------------------------------------------------
    def _ws(self, solve: bool):
        import torch

        from . import hip

        lib = hip.require_gpu()
        nbytes = int(lib.cc_sumfact_workspace_bytes(ctypes.addressof(self.desc), 1 if solve else 0))
        if nbytes == 0:
            raise ValueError("invalid matrix-free descriptor")
        key = "ws_solve" if solve else "ws"
        buf = getattr(self, key, None)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            setattr(self, key, buf)
        return lib, buf, nbytes

    def apply(self, x, y=None):
        """``y = E^T (K + c M) E x`` on the device (float64)."""
        import torch

        from . import hip

        lib, ws, nbytes = self._ws(False)
        y = torch.empty_like(x) if y is None else y
        hip._check(lib.cc_sumfact_apply(ctypes.addressof(self.desc), hip._ptr(x), hip._ptr(y), hip._ptr(ws), nbytes,
                                        hip._stream()), "sumfact_apply")
        return y

    def pcg(self, b, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi", check_every=16, history=False, x=None,
            snapshots=None):
        """Device Jacobi-PCG on the matrix-free operator. Returns ``(x, result[, history][, snaps])``;
        ``snapshots`` (iteration numbers) adds a ``[len, n]`` device tensor of
        the iterates after those iterations (the final one where the solve
        stopped earlier)."""
        import torch

        from . import hip

        lib, ws, nbytes = self._ws(True)
        x = torch.empty_like(b) if x is None else x
        res = hip.PcgResult()
        hist = torch.full((int(maxiter) + 1,), float("nan"), dtype=torch.float64, device=b.device) if history else None
        n_snap = 0 if snapshots is None else len(snapshots)
        snap_iters = (ctypes.c_int32 * max(n_snap, 1))(*([int(k) for k in snapshots] if n_snap else [0]))
        snaps = torch.empty((n_snap, self.n), dtype=torch.float64, device=b.device) if n_snap else None
        hip._check(lib.cc_pcg_sumfact(ctypes.addressof(self.desc), hip._ptr(b), hip._ptr(x), float(tol), int(maxiter),
                                      {"none": 0, "jacobi": 1}[preconditioner], int(check_every), hip._ptr(ws), nbytes,
                                      ctypes.addressof(res), hip._ptr(hist) if hist is not None else None,
                                      snap_iters, n_snap, hip._ptr(snaps) if snaps is not None else None,
                                      hip._stream()), "pcg_sumfact")
        out = (x, res)
        if history:
            out += (hist.cpu().numpy(),)
        if n_snap:
            out += (snaps,)
        return out

    # -- k right-hand sides at once -----------------------------------------------
    def _ws_k(self, k: int, solve: bool):
        import torch

        from . import hip

        lib = hip.require_gpu()
        nbytes = int(lib.cc_sumfact_k_workspace_bytes(ctypes.addressof(self.desc), int(k), 1 if solve else 0))
        if nbytes == 0:
            raise ValueError(f"invalid matrix-free descriptor or k={k} (k in {hip.SUPPORTED_K})")
        key = f"ws_k{k}_{int(solve)}"
        buf = getattr(self, key, None)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            setattr(self, key, buf)
        return lib, buf, nbytes

    def apply_k(self, X, Y=None):
        """``Y = E^T (K + c M) E X`` for a row-major ``[n, k]`` float64 block."""
        import torch

        from . import hip

        k = X.shape[1]
        lib, ws, nbytes = self._ws_k(k, False)
        X = X.contiguous()
        Y = torch.empty_like(X) if Y is None else Y
        hip._check(lib.cc_sumfact_apply_k(ctypes.addressof(self.desc), int(k), hip._ptr(X), hip._ptr(Y), hip._ptr(ws),
                                          nbytes, hip._stream()), "sumfact_apply_k")
        return Y

    @property
    def blocks32(self) -> int:
        """Device address of the float32 copy of the cut-cell corrections (made
        on first use) for the float32 operator of the mixed-precision PCG."""
        import torch

        if self._blocks32 is None:
            self._blocks32 = self._blocks64.to(torch.float32)
        return self._blocks32.data_ptr()

    def apply_k_f32(self, X, Y=None):
        """``Y = E^T (K + c M) E X`` in float32 (the inner operator of
        ``pcg_mixed_batched``) for a row-major ``[n, k]`` float32 block."""
        import torch

        from . import hip

        k = X.shape[1]
        if X.dtype != torch.float32:
            raise TypeError("apply_k_f32 takes a float32 block")
        lib, ws, nbytes = self._ws_k(k, False)
        X = X.contiguous()
        Y = torch.empty_like(X) if Y is None else Y
        hip._check(lib.cc_sumfact_apply_k_f32(ctypes.addressof(self.desc), self.blocks32, int(k), hip._ptr(X),
                                              hip._ptr(Y), hip._ptr(ws), nbytes, hip._stream()),
                   "sumfact_apply_k_f32")
        return Y

    def pcg_mixed_batched(self, B, *, tol=1.0e-8, maxiter=5000, inner_tol=1.0e-4, max_outer=12,
                          preconditioner="jacobi", check_every=16, X=None):
        """FP32 inner / FP64 outer Jacobi-PCG on the ``k`` columns of ``B``:
        the inner solves apply the operator in float32 to ``inner_tol``, the
        outer loop stops on the float64 residual ``||B - A X|| / ||B|| <= tol``.
        Returns ``(X, [result per column])``; ``iterations`` counts inner
        iterations and ``outer`` the sweeps."""
        import torch

        from . import hip

        k = B.shape[1]
        lib = hip.require_gpu()
        nbytes = int(lib.cc_sumfact_mixed_k_workspace_bytes(ctypes.addressof(self.desc), int(k)))
        if nbytes == 0:
            raise ValueError(f"invalid matrix-free descriptor or k={k} (k in {hip.SUPPORTED_K})")
        key = f"ws_mixed_k{k}"
        ws = getattr(self, key, None)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            setattr(self, key, ws)
        B = B.contiguous()
        X = torch.empty_like(B) if X is None else X
        res = (hip.PcgResult * k)()
        hip._check(lib.cc_pcg_sumfact_mixed_k(ctypes.addressof(self.desc), self.blocks32, int(k), hip._ptr(B),
                                              hip._ptr(X), float(tol), int(maxiter), float(inner_tol),
                                              int(max_outer), {"none": 0, "jacobi": 1}[preconditioner],
                                              int(check_every), hip._ptr(ws), nbytes, res, hip._stream()),
                   "pcg_sumfact_mixed_k")
        return X, list(res)

    def pcg_batched(self, B, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi", check_every=16, X=None):
        """Jacobi-PCG on the ``k`` columns of ``B`` (``[n, k]``, k in {1, 2, 4,
        8, 16}) in one device loop, each cut block read once per apply for all
        columns. Returns ``(X, [result per column])``."""
        import torch

        from . import hip

        k = B.shape[1]
        lib, ws, nbytes = self._ws_k(k, True)
        B = B.contiguous()
        X = torch.empty_like(B) if X is None else X
        res = (hip.PcgResult * k)()
        hip._check(lib.cc_pcg_sumfact_k(ctypes.addressof(self.desc), int(k), hip._ptr(B), hip._ptr(X), float(tol),
                                        int(maxiter), {"none": 0, "jacobi": 1}[preconditioner], int(check_every),
                                        hip._ptr(ws), nbytes, res, hip._stream()), "pcg_sumfact_k")
        return X, list(res)
