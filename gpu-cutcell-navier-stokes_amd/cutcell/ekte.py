"""The aggregated operator ``E^T B E`` applied without forming it.

The reference forms ``E^T A E`` by a sparse triple product (``restrict_spd``,
``code/bell_3d_core.py:550-552``) and multiplies by it
(``code/fast_solver_backends.py:250``). ``EkteParts`` keeps the factors
instead, renumbered so that the HIP product (``cc_ekte_apply``,
``csrc/ekte.hip``) reads the reduced unknowns directly:

- ``B``: the background operator (``K + H``, symmetrised) over the active
  dofs; row/column ``j < n_red`` is the interior dof of reduced unknown ``j``,
  ``n_red + s`` is small dof ``s``;
- ``E_s``: the small dofs' extension rows of ``E`` (over reduced unknowns);
- ``T``: ``E_s^T`` restricted to the reduced unknowns it reaches;
- ``diag``: the diagonal of ``E^T B E``, computed from the factors (no triple
  product): ``B_jj + 2 sum_s E_sj B_sj + sum_st E_sj B_st E_tj``.

The set-up is host work on SciPy sparse matrices; the apply and the solve run
on the GPU only.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import hip


@dataclass
class EkteParts:
    n_red: int
    n_small: int
    B: sp.csr_matrix  # (n_red + n_small)^2, combined numbering
    Es: sp.csr_matrix  # (n_small, n_red)
    t_rows: np.ndarray  # reduced unknowns reached by some small dof
    T: sp.csr_matrix  # (n_touch, n_small): row i is column t_rows[i] of Es
    diag: np.ndarray

    @property
    def nnz_b(self) -> int:
        return int(self.B.nnz)

    def algorithmic_bytes(self, k: int = 1) -> int:
        """Bytes one apply reads and writes once each: B (values and
        indices), row pointers, E_s and T (values and indices), the vectors
        x and y, and the small dofs' u_s and v_s (written and read)."""
        nb = self.n_red + self.n_small
        b = self.B.nnz * 12 + 4 * (nb + 1)
        b += (self.Es.nnz + self.T.nnz) * 12 + 4 * (self.n_small + 1) + 4 * (2 * self.t_rows.size + 1)
        b += 2 * self.n_red * k * 8 + 4 * self.n_small * k * 8
        return int(b)

    def to_scipy_operator(self) -> sp.csr_matrix:
        """``E^T B E`` formed on the host from the parts (for tests)."""
        n = self.n_red
        E = sp.vstack([sp.identity(n, format="csr"), self.Es]).tocsr()
        return (E.T @ self.B @ E).tocsr()


def ekte_parts(B_full, E, interior, small) -> EkteParts:
    """Parts of ``E^T B E`` from the background operator ``B_full``
    (``n_full^2``), the extension ``E`` (``n_full x n_red``), the interior
    dofs in reduced order and the small dofs."""
    B_full = sp.csr_matrix(B_full)
    B_full = (0.5 * (B_full + B_full.T)).tocsr()
    E = sp.csr_matrix(E)
    interior = np.asarray(interior, dtype=np.int64)
    small = np.asarray(small, dtype=np.int64)
    n_red, n_small = interior.size, small.size
    order = np.concatenate([interior, small])
    B = B_full[order][:, order].tocsr()
    B.sort_indices()
    Es = E[small].tocsr()
    Es.sort_indices()
    EsT = Es.T.tocsr()  # (n_red, n_small)
    EsT.sort_indices()
    t_rows = np.nonzero(np.diff(EsT.indptr))[0]
    T = EsT[t_rows].tocsr()
    # diag(E^T B E) from the factors.
    d = B.diagonal()[:n_red].copy()
    if n_small:
        B_si = B[n_red:, :n_red]
        B_ss = B[n_red:, n_red:]
        d += 2.0 * np.asarray(Es.multiply(B_si).sum(axis=0)).ravel()
        d += np.asarray(Es.multiply(B_ss @ Es).sum(axis=0)).ravel()
    return EkteParts(n_red, n_small, B, Es, t_rows, T, d)


def ekte_parts_3d(levelset, n_elements: int, degree: int, *, quadrature_order: int | None = None,
                  support_threshold: float = 0.04, tie_order: str = "stable"):
    """Parts of the 3D cut-box operator ``E^T (K + H) E`` (the system
    ``cutcell.cut3d.build_system_3d`` forms), with no triple product."""
    from .cut3d import assemble_levelset_3d, polynomial_preserving_extension_3d

    asm = assemble_levelset_3d(levelset, n_elements, degree, quadrature_order)
    agg = polynomial_preserving_extension_3d(asm, support_threshold=support_threshold, tie_order=tie_order)
    return ekte_parts(asm.K + asm.H, agg.E, agg.interior, agg.small_support), asm, agg


class EkteDesc(ctypes.Structure):
    """``cc_ekte_desc`` of the C header."""

    _fields_ = [
        ("n_red", ctypes.c_int32),
        ("n_small", ctypes.c_int32),
        ("n_touch", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("nnz_b", ctypes.c_int64),
        ("b_indptr", ctypes.c_void_p),
        ("b_indices", ctypes.c_void_p),
        ("b_vals", ctypes.c_void_p),
        ("e_indptr", ctypes.c_void_p),
        ("e_indices", ctypes.c_void_p),
        ("e_vals", ctypes.c_void_p),
        ("t_rows", ctypes.c_void_p),
        ("t_indptr", ctypes.c_void_p),
        ("t_small", ctypes.c_void_p),
        ("t_vals", ctypes.c_void_p),
        ("b_rb", ctypes.c_void_p),
        ("diag", ctypes.c_void_p),
    ]


class DeviceEkte:
    """Device arrays of ``EkteParts`` and the packed descriptor. With
    ``row_blocks=True`` the product on ``B`` takes the row-block form
    (groups of 4 rows sharing each gathered x row, built on the device)."""

    def __init__(self, parts: EkteParts, device: str = "cuda", row_blocks: bool = False):
        import torch

        self.parts = parts
        keep = []

        def t(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            x = torch.from_numpy(a if a.size else np.zeros(1, dtype=dt)).to(device)
            keep.append(x)
            return x.data_ptr()

        d = EkteDesc()
        d.n_red, d.n_small, d.n_touch = parts.n_red, parts.n_small, int(parts.t_rows.size)
        d.nnz_b = parts.B.nnz
        d.b_indptr = t(parts.B.indptr, np.int32)
        d.b_indices = t(parts.B.indices, np.int32)
        d.b_vals = t(parts.B.data, np.float64)
        self.rb = None
        if row_blocks:
            nb = parts.n_red + parts.n_small
            Bd = hip.DeviceCSR(nb, parts.B.nnz, keep[0], keep[1], keep[2])
            self.rb = hip.DeviceRowBlocks.from_csr(Bd, fp64=True, fp32=False)
            d.b_rb = ctypes.addressof(self.rb.desc)
        d.e_indptr = t(parts.Es.indptr, np.int32)
        d.e_indices = t(parts.Es.indices, np.int32)
        d.e_vals = t(parts.Es.data, np.float64)
        d.t_rows = t(parts.t_rows, np.int32)
        d.t_indptr = t(parts.T.indptr, np.int32)
        d.t_small = t(parts.T.indices, np.int32)
        d.t_vals = t(parts.T.data, np.float64)
        d.diag = t(parts.diag, np.float64)
        self.desc = d
        self._keep = keep
        self.n = parts.n_red
        self.device = device
        self._ws = hip.Workspace()

    def _workspace(self, k: int, solve: int):
        lib = hip.require_gpu()
        nbytes = int(lib.cc_ekte_workspace_bytes(ctypes.addressof(self.desc), k, solve))
        if nbytes == 0:
            raise hip.HipError("cc_ekte_workspace_bytes rejected the descriptor or k")
        return self._ws.get(nbytes, self._keep[0].device), nbytes

    def apply(self, X, Y=None):
        """``Y = E^T B E X`` for ``X`` of shape ``[n]`` or ``[n, k]`` (float64)."""
        import torch

        lib = hip.require_gpu()
        X2 = X.reshape(self.n, -1).contiguous()
        k = X2.shape[1]
        if k not in hip.SUPPORTED_K:
            raise ValueError(f"k must be one of {hip.SUPPORTED_K}")
        Y2 = torch.empty_like(X2) if Y is None else Y.reshape(self.n, k)
        ws, nbytes = self._workspace(k, 0)
        hip._check(lib.cc_ekte_apply(ctypes.addressof(self.desc), k, hip._ptr(X2), hip._ptr(Y2), hip._ptr(ws),
                                     nbytes, hip._stream()), "ekte_apply")
        return Y2.reshape(X.shape)

    def pcg(self, B, *, tol: float = 1.0e-8, maxiter: int = 5000, preconditioner: str = "jacobi",
            check_every: int = 16):
        """Jacobi-PCG on the fused operator for ``B`` of shape ``[n]`` or
        ``[n, k]``; returns ``(X, results)``."""
        import torch

        lib = hip.require_gpu()
        B2 = B.reshape(self.n, -1).to(torch.float64).contiguous()
        k = B2.shape[1]
        if k not in hip.SUPPORTED_K:
            raise ValueError(f"k must be one of {hip.SUPPORTED_K}")
        X2 = torch.empty_like(B2)
        ws, nbytes = self._workspace(k, 1)
        res = (hip.PcgResult * k)()
        prec = {"none": 0, "jacobi": 1}[preconditioner]
        hip._check(lib.cc_pcg_ekte(ctypes.addressof(self.desc), k, hip._ptr(B2), hip._ptr(X2), float(tol),
                                   int(maxiter), prec, int(check_every), hip._ptr(ws), nbytes,
                                   ctypes.addressof(res), hip._stream()), "pcg_ekte")
        return (X2.reshape(-1) if B.dim() == 1 else X2), list(res)
