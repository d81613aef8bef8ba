"""ctypes binding of the gfx950 HIP library (``libcutcell_hip.so``).

PyTorch-ROCm is used only to own device memory and to name the current HIP
stream; every computation runs in the hand-written kernels behind the C ABI
declared in ``include/cutcell_hip.h``. There is no CPU code path here: when
the library or a GPU is missing the calls raise ``HipUnavailable``.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcutcell_hip.so"

CC_RUNNING, CC_CONVERGED, CC_BREAKDOWN, CC_MAXITER = 0, 1, 2, 3
STATUS_MESSAGE = {
    CC_CONVERGED: "converged",
    CC_BREAKDOWN: "CG breakdown: non-positive curvature",
    CC_MAXITER: "maximum iterations reached",
    CC_RUNNING: "maximum iterations reached",
}
SUPPORTED_K = (1, 2, 4, 8, 16)


class HipUnavailable(RuntimeError):
    """The HIP library or a GPU is not available."""


class HipError(RuntimeError):
    """A library call returned a nonzero status."""


class PcgResult(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("outer", ctypes.c_int32),
        ("inner_total", ctypes.c_int32),
        ("recursive_rel", ctypes.c_double),
        ("true_rel", ctypes.c_double),
    ]


class GmresResult(ctypes.Structure):
    """``cc_gmres_result`` of the C header."""

    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("info", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("residual_norm", ctypes.c_double),
        ("b_norm", ctypes.c_double),
    ]


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_F64 = ctypes.c_double
_SZ = ctypes.c_size_t

_SIGNATURES = {
    "cc_version": ([], ctypes.c_int),
    "cc_device_count": ([_P], ctypes.c_int),
    "cc_profile_enable": ([_I32], ctypes.c_int),
    "cc_profile_enable_classes": ([_I32], ctypes.c_int),
    "cc_profile_reset": ([], ctypes.c_int),
    "cc_profile_read": ([_I32, _P, _P], ctypes.c_int),
    "cc_profile_collect": ([], ctypes.c_int),
    "cc_spmm_set_config": ([_I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32], ctypes.c_int),
    "cc_spmm_csr_f32": ([_I32, _I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_sumfact_workspace_bytes": ([_P, _I32], _SZ),
    "cc_sumfact_apply": ([_P, _P, _P, _P, _SZ, _P], ctypes.c_int),
    "cc_pcg_sumfact": ([_P, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P, _P, _I32, _P, _P], ctypes.c_int),
    "cc_sumfact_k_workspace_bytes": ([_P, _I32, _I32], _SZ),
    "cc_sumfact_apply_k": ([_P, _I32, _P, _P, _P, _SZ, _P], ctypes.c_int),
    "cc_pcg_sumfact_k": ([_P, _I32, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P], ctypes.c_int),
    "cc_cut3d_gram": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "cc_sumfact_mixed_k_workspace_bytes": ([_P, _I32], _SZ),
    "cc_sumfact_apply_k_f32": ([_P, _P, _I32, _P, _P, _P, _SZ, _P], ctypes.c_int),
    "cc_pcg_sumfact_mixed_k": ([_P, _P, _I32, _P, _P, _F64, _I32, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P],
                               ctypes.c_int),
    "cc_pcg_batched": (
        [_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _F64, _I32, _I32, _I32,
         _I32, _P, _P, _P],
        ctypes.c_int,
    ),
    "cc_spmv_csr_f64": ([_I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_spmv_csr_f32": ([_I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_spmm_csr_f64": ([_I32, _I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_jacobi_diag_f64": ([_I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_residual_f64": ([_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _SZ, _P], ctypes.c_int),
    "cc_pcg_workspace_bytes": ([_I32, _I32, _I32], _SZ),
    "cc_pcg_jacobi_f64": ([_I32, _I32, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P, _P], ctypes.c_int),
    "cc_pcg_noalloc_max_bytes": ([], _SZ),
    "cc_pcg_jacobi_f64_noalloc": ([_I32, _I32, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _P, _P, _P], ctypes.c_int),
    "cc_pcg_jacobi_mixed": (
        [_I32, _I32, _P, _P, _P, _P, _P, _P, _F64, _I32, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P],
        ctypes.c_int,
    ),
    "cc_ekte_workspace_bytes": ([_P, _I32, _I32], _SZ),
    "cc_ekte_apply": ([_P, _I32, _P, _P, _P, _SZ, _P], ctypes.c_int),
    "cc_pcg_ekte": ([_P, _I32, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P], ctypes.c_int),
    "cc_gmres_workspace_bytes": ([_I32, _I32], _SZ),
    "cc_gmres_jacobi_f64": ([_I32, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P], ctypes.c_int),
    "cc_cut2d_gram": ([_I32, _P, _I32, _I32, _I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_asm2d_count": ([_I32, _I32, _P, _P, _P, _P], ctypes.c_int),
    "cc_asm2d_fill": ([_I32, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_rb_build_count": ([_I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_rb_build_fill": ([_I32, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_rb_build_pairs_count": ([_I32, _P, _P, _I32, _P, _P, _P, _P], ctypes.c_int),
    "cc_rb_build_pairs_fill": ([_I32, _P, _P, _P, _I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_rb_set_config": ([_I32, _I32, _I32, _I32, _I32, _I32], ctypes.c_int),
    "cc_spmm_rb_f64": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "cc_spmm_rb_f32": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "cc_runs_count": ([_I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_runs_fill": ([_I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_spmv_runs_f64": ([_I32, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_spmv_runs_f32": ([_I32, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "cc_pcg_jacobi_f64_runs": (
        [_I32, _P, _P, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P, _P], ctypes.c_int,
    ),
    "cc_pcg_jacobi_mixed_runs": (
        [_I32, _P, _P, _P, _P, _P, _P, _P, _P, _F64, _I32, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P], ctypes.c_int,
    ),
    "cc_pcg_jacobi_f64_rb": (
        [_I32, _I32, _P, _P, _P, _P, _P, _P, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P, _P], ctypes.c_int,
    ),
    "cc_pcg_jacobi_mixed_rb": (
        [_I32, _I32, _P, _P, _P, _P, _P, _P, _F64, _I32, _F64, _I32, _I32, _I32, _P, _SZ, _P, _P],
        ctypes.c_int,
    ),
}

RB_ROWS = 4  # CC_RB_ROWS of the C header


class RbDesc(ctypes.Structure):
    """``cc_rb_desc`` of the C header."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("ngroups", ctypes.c_int32),
        ("gptr", ctypes.c_void_p),
        ("gcol", ctypes.c_void_p),
        ("gval", ctypes.c_void_p),
        ("gval32", ctypes.c_void_p),
        ("grptr", ctypes.c_void_p),
        ("grun", ctypes.c_void_p),
        ("pairs", ctypes.c_void_p),
    ]


def declared_symbols() -> list[str]:
    """Every entry point the Python side binds (all declared in the C header)."""
    return sorted(_SIGNATURES)


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load the library and declare its signatures (no GPU needed to load)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise HipUnavailable(f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _torch():
    import torch

    return torch


def require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise HipUnavailable("no ROCm GPU is visible to this process")
    return load_library()


CC_E_TOO_LARGE = -4


def _check(status: int, what: str) -> None:
    if status == CC_E_TOO_LARGE:
        raise HipError(f"{what}: the problem needs more device scratch than the library's fixed array holds "
                       f"(status {status})")
    if status != 0:
        raise HipError(f"{what} failed with status {status}")


def _ptr(t) -> int:
    return int(t.data_ptr())


def _stream() -> int:
    return int(_torch().cuda.current_stream().cuda_stream)


@dataclass
class DeviceCSR:
    """A CSR matrix resident in HBM (int32 index arrays, float64 values)."""

    n: int
    nnz: int
    indptr: object
    indices: object
    vals: object
    vals32: object | None = None

    @staticmethod
    def from_scipy(A, device: str = "cuda", with_f32: bool = False) -> "DeviceCSR":
        torch = _torch()
        A = A.tocsr()
        if A.shape[0] != A.shape[1]:
            raise ValueError("a square matrix is required")
        if A.nnz >= 2**31:
            raise ValueError("nnz must fit in int32")
        indptr = torch.from_numpy(np.ascontiguousarray(A.indptr, dtype=np.int32)).to(device)
        indices = torch.from_numpy(np.ascontiguousarray(A.indices, dtype=np.int32)).to(device)
        vals = torch.from_numpy(np.ascontiguousarray(A.data, dtype=np.float64)).to(device)
        d = DeviceCSR(int(A.shape[0]), int(A.nnz), indptr, indices, vals)
        if with_f32:
            d.ensure_f32()
        return d

    def ensure_f32(self):
        if self.vals32 is None:
            self.vals32 = self.vals.to(_torch().float32)
        return self.vals32

    def bytes(self, value_bytes: int = 8) -> int:
        return int(self.nnz * (value_bytes + 4) + 4 * (self.n + 1))


@dataclass
class DeviceRowBlocks:
    """Row-block form of a device CSR matrix (``cc_rb_desc``): groups of
    ``RB_ROWS`` consecutive rows, the ascending union of their columns and,
    per union column, the group's row values (zero where a row has none).
    Built on the device from ``A`` (column indices must ascend in each row)."""

    n: int
    ngroups: int
    nu: int
    gptr: object
    gcol: object
    gval: object | None
    gval32: object | None
    desc: RbDesc
    grptr: object | None = None
    grun: object | None = None
    nruns: int = 0
    pairs: object | None = None

    @staticmethod
    def from_csr(A: DeviceCSR, fp64: bool = True, fp32: bool = True, runs: bool = False) -> "DeviceRowBlocks":
        """``runs``: also encode each group's union columns as runs of
        consecutive columns (``grptr``, ``grun``); products whose group team
        is one wave then read the columns from the runs."""
        lib = require_gpu()
        torch = _torch()
        dev = A.vals.device
        ng = -(-A.n // RB_ROWS)
        gptr = torch.empty(ng + 1, dtype=torch.int32, device=dev)
        nu = ctypes.c_int64()
        _check(lib.cc_rb_build_count(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(gptr), ctypes.addressof(nu),
                                     _stream()), "rb_build_count (column indices must ascend within rows)")
        nu = int(nu.value)
        gcol = torch.empty(max(nu, 1), dtype=torch.int32, device=dev)
        gval = torch.empty(max(nu, 1) * RB_ROWS, dtype=torch.float64, device=dev) if fp64 else None
        gval32 = torch.empty(max(nu, 1) * RB_ROWS, dtype=torch.float32, device=dev) if fp32 else None
        _check(lib.cc_rb_build_fill(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(gptr), _ptr(gcol),
                                    _ptr(gval) if fp64 else None, _ptr(gval32) if fp32 else None, _stream()),
               "rb_build_fill")
        desc = RbDesc(A.n, ng, _ptr(gptr), _ptr(gcol), _ptr(gval) if fp64 else None,
                      _ptr(gval32) if fp32 else None)
        out = DeviceRowBlocks(A.n, ng, nu, gptr, gcol, gval, gval32, desc)
        if runs:
            grptr = torch.empty(ng + 1, dtype=torch.int32, device=dev)
            nr = ctypes.c_int64()
            _check(lib.cc_runs_count(ng, _ptr(gptr), _ptr(gcol), _ptr(grptr), ctypes.addressof(nr), _stream()),
                   "runs_count (row-block union columns)")
            nr = int(nr.value)
            grun = torch.empty(2 * max(nr, 1), dtype=torch.int32, device=dev)
            _check(lib.cc_runs_fill(ng, _ptr(gptr), _ptr(gcol), _ptr(grptr), _ptr(grun), _stream()), "runs_fill")
            out.grptr, out.grun, out.nruns = grptr, grun, nr
            desc.grptr, desc.grun = _ptr(grptr), _ptr(grun)
        return out

    @staticmethod
    def from_csr_pairs(A: DeviceCSR, pairs: np.ndarray, fp64: bool = True, fp32: bool = True) -> "DeviceRowBlocks":
        """Paired row blocks: the 8 rows of the two 4-row groups of each row
        of ``pairs`` (``cutcell.rowpairs.row_group_pairs``; every group once)
        form one group, with 8 values per union column."""
        lib = require_gpu()
        torch = _torch()
        dev = A.vals.device
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        ng4 = -(-A.n // RB_ROWS)
        seen = np.concatenate([pairs[:, 0], pairs[pairs[:, 1] >= 0, 1]])
        if seen.size != ng4 or not np.array_equal(np.sort(seen), np.arange(ng4)):
            raise ValueError("pairs must list every 4-row group exactly once")
        npairs = int(pairs.shape[0])
        pd = torch.from_numpy(pairs.reshape(-1)).to(dev)
        gptr = torch.empty(npairs + 1, dtype=torch.int32, device=dev)
        nu = ctypes.c_int64()
        _check(lib.cc_rb_build_pairs_count(A.n, _ptr(A.indptr), _ptr(A.indices), npairs, _ptr(pd), _ptr(gptr),
                                           ctypes.addressof(nu), _stream()),
               "rb_build_pairs_count (column indices must ascend within rows)")
        nu = int(nu.value)
        R2 = 2 * RB_ROWS
        gcol = torch.empty(max(nu, 1), dtype=torch.int32, device=dev)
        gval = torch.empty(max(nu, 1) * R2, dtype=torch.float64, device=dev) if fp64 else None
        gval32 = torch.empty(max(nu, 1) * R2, dtype=torch.float32, device=dev) if fp32 else None
        _check(lib.cc_rb_build_pairs_fill(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), npairs, _ptr(pd),
                                          _ptr(gptr), _ptr(gcol), _ptr(gval) if fp64 else None,
                                          _ptr(gval32) if fp32 else None, _stream()), "rb_build_pairs_fill")
        desc = RbDesc(A.n, npairs, _ptr(gptr), _ptr(gcol), _ptr(gval) if fp64 else None,
                      _ptr(gval32) if fp32 else None, None, None, _ptr(pd))
        return DeviceRowBlocks(A.n, npairs, nu, gptr, gcol, gval, gval32, desc, pairs=pd)

    def bytes(self, value_bytes: int = 8) -> int:
        """Stored bytes one product reads: group offsets (and the pair list),
        the union columns (or, with runs, the run offsets and runs) and the
        padded values of one precision."""
        rows = 2 * RB_ROWS if self.pairs is not None else RB_ROWS
        cols = 4 * (self.ngroups + 1) + 8 * self.nruns if self.grun is not None else 4 * self.nu
        extra = 8 * self.ngroups if self.pairs is not None else 0
        return int(4 * (self.ngroups + 1) + cols + extra + self.nu * rows * value_bytes)


@dataclass
class DeviceRuns:
    """Run encoding of a device CSR matrix's rows (``cc_runs_count`` /
    ``cc_runs_fill``): per row its runs of consecutive column indices, 8
    bytes each (first column; offset within the row << 8 | length), in place
    of one 4-byte index per nonzero. The values stay the CSR arrays."""

    n: int
    nruns: int
    rptr: object
    runs: object
    csr: DeviceCSR

    @staticmethod
    def from_csr(A: DeviceCSR) -> "DeviceRuns":
        lib = require_gpu()
        torch = _torch()
        dev = A.vals.device
        rptr = torch.empty(A.n + 1, dtype=torch.int32, device=dev)
        nr = ctypes.c_int64()
        _check(lib.cc_runs_count(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(rptr), ctypes.addressof(nr), _stream()),
               "runs_count (column indices must ascend strictly within rows)")
        nr = int(nr.value)
        runs = torch.empty(2 * max(nr, 1), dtype=torch.int32, device=dev)
        _check(lib.cc_runs_fill(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(rptr), _ptr(runs), _stream()), "runs_fill")
        return DeviceRuns(A.n, nr, rptr, runs, A)

    def bytes(self, value_bytes: int = 8) -> int:
        """Bytes one product reads from the matrix: the values, the runs and
        both row pointers."""
        return int(self.csr.nnz * value_bytes + 8 * self.nruns + 8 * (self.n + 1))


def spmv_runs(R: DeviceRuns, x, y=None):
    """y = A x from the run encoding (float64 or float32 by x's dtype)."""
    lib = require_gpu()
    torch = _torch()
    if y is None:
        y = torch.empty_like(x)
    if x.dtype == torch.float64:
        _check(lib.cc_spmv_runs_f64(R.n, _ptr(R.rptr), _ptr(R.runs), _ptr(R.csr.indptr), _ptr(R.csr.vals), _ptr(x),
                                    _ptr(y), _stream()), "spmv_runs_f64")
    else:
        _check(lib.cc_spmv_runs_f32(R.n, _ptr(R.rptr), _ptr(R.runs), _ptr(R.csr.indptr), _ptr(R.csr.ensure_f32()),
                                    _ptr(x), _ptr(y), _stream()), "spmv_runs_f32")
    return y


def spmm_rb(rb: DeviceRowBlocks, X, Y=None):
    """``Y = A X`` from the row-block form (``X`` float64 or float32, ``[n, k]``)."""
    lib = require_gpu()
    torch = _torch()
    X2 = X.reshape(rb.n, -1).contiguous()
    k = X2.shape[1]
    if k not in SUPPORTED_K:
        raise ValueError(f"k must be one of {SUPPORTED_K}")
    Y = torch.empty_like(X2) if Y is None else Y
    fn = lib.cc_spmm_rb_f64 if X2.dtype == torch.float64 else lib.cc_spmm_rb_f32
    _check(fn(ctypes.addressof(rb.desc), k, _ptr(X2), _ptr(Y), _stream()), "spmm_rb")
    return Y.reshape(X.shape)


def spmv(A: DeviceCSR, x, y=None):
    lib = require_gpu()
    torch = _torch()
    if x.dtype == torch.float64:
        y = torch.empty_like(x) if y is None else y
        _check(lib.cc_spmv_csr_f64(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(x), _ptr(y), _stream()), "spmv")
    elif x.dtype == torch.float32:
        y = torch.empty_like(x) if y is None else y
        v32 = A.ensure_f32()
        _check(lib.cc_spmv_csr_f32(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(v32), _ptr(x), _ptr(y), _stream()), "spmv")
    else:
        raise TypeError("x must be float64 or float32")
    return y


def spmm(A: DeviceCSR, X, Y=None):
    """``Y = A X`` for a row-major ``[n, k]`` float64 (or float32, with the
    float32 copy of the values) block, k in {1,2,4,8,16}."""
    lib = require_gpu()
    torch = _torch()
    k = X.shape[1]
    if k not in SUPPORTED_K:
        raise ValueError(f"k must be one of {SUPPORTED_K}")
    X = X.contiguous()
    Y = torch.empty_like(X) if Y is None else Y
    if X.dtype == torch.float64:
        fn, vals = lib.cc_spmm_csr_f64, A.vals
    elif X.dtype == torch.float32:
        fn, vals = lib.cc_spmm_csr_f32, A.ensure_f32()
    else:
        raise TypeError("X must be float64 or float32")
    _check(fn(A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(vals), _ptr(X), _ptr(Y), _stream()), "spmm")
    return Y


def jacobi_diag(A: DeviceCSR):
    lib = require_gpu()
    d = _torch().empty(A.n, dtype=_torch().float64, device=A.vals.device)
    _check(lib.cc_jacobi_diag_f64(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(d), _stream()), "diag")
    return d


def residual(A: DeviceCSR, B, X) -> np.ndarray:
    """FP64 ``||B - A X|| / ||B||`` per column (device computation)."""
    lib = require_gpu()
    torch = _torch()
    B2 = B.reshape(A.n, -1).contiguous()
    X2 = X.reshape(A.n, -1).contiguous()
    k = B2.shape[1]
    if k not in SUPPORTED_K:
        raise ValueError(f"k must be one of {SUPPORTED_K}")
    ws_bytes = A.n * k * 8 + 1024 * 2 * k * 8 + 4096
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=A.vals.device)
    rel = np.zeros(k)
    _check(
        lib.cc_residual_f64(
            A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(B2), _ptr(X2),
            rel.ctypes.data, _ptr(ws), ws_bytes, _stream(),
        ),
        "residual",
    )
    return rel


PROFILE_CLASSES = {"spmm": 0, "update": 1, "direction": 2, "batched": 3, "learned": 4}


def profile_enable(on=True) -> None:
    """Event timing of the solver kernels: True (every class), False (off) or
    an iterable of class names (only those: each timed launch adds its two
    events' dispatch to the stream)."""
    lib = load_library()
    if isinstance(on, bool):
        lib.cc_profile_enable(1 if on else 0)
    else:
        mask = 0
        for name in on:
            mask |= 1 << PROFILE_CLASSES[name]
        lib.cc_profile_enable_classes(mask)
    lib.cc_profile_reset()


def profile_read(which: str) -> tuple[float, int]:
    """(summed milliseconds, launch count) of a timed kernel class.

    Synchronises the device first so that launches made outside the solvers'
    own polling loops are counted."""
    lib = load_library()
    _torch().cuda.synchronize()
    lib.cc_profile_collect()
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    _check(lib.cc_profile_read(PROFILE_CLASSES[which], ctypes.addressof(ms), ctypes.addressof(cnt)), "profile_read")
    return float(ms.value), int(cnt.value)


class Workspace:
    """A reusable device workspace grown on demand."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes: int, device):
        torch = _torch()
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != device:
            self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        return self.buf


_default_ws = Workspace()


def pcg(
    A: DeviceCSR,
    B,
    *,
    tol: float = 1.0e-8,
    maxiter: int = 5000,
    preconditioner: str = "jacobi",
    precision: str = "fp64",
    inner_tol: float = 1.0e-4,
    max_outer: int = 12,
    check_every: int = 16,
    X=None,
    workspace: Workspace | None = None,
    history: bool = False,
    tune: bool = False,
    rb: DeviceRowBlocks | None = None,
    runs: "DeviceRuns | None" = None,
):
    """Device Jacobi-PCG on ``k`` right-hand sides (``B`` is ``[n]`` or ``[n, k]``).

    Returns ``(X, results)`` where ``results`` is a list of ``PcgResult``
    (and, with ``history=True`` for ``precision="fp64"``, a ``[maxiter+1, k]``
    array of ``||r_j||/||b||`` per iteration, NaN past each column's stop).
    ``precision="mixed"`` runs float32 inner PCG under float64 refinement.
    With ``rb`` (the row-block form of ``A``) the CG products use it; with
    ``runs`` (the run encoding of ``A``'s rows, one right-hand side only) they
    come from the runs.
    """
    lib = require_gpu()
    torch = _torch()
    squeeze = B.dim() == 1
    B2 = B.reshape(A.n, -1).to(torch.float64).contiguous()
    k = B2.shape[1]
    if k not in SUPPORTED_K:
        raise ValueError(f"k must be one of {SUPPORTED_K}")
    X2 = torch.empty_like(B2) if X is None else X.reshape(A.n, k)
    prec = {"none": 0, "jacobi": 1}[preconditioner]
    mode = {"fp64": 0, "mixed": 1}[precision]
    if tune:
        from .autotune import rb_config, spmm_config

        spmm_config(A, k, precision)
        if rb is not None:
            rb_config(rb, k, precision)
    nbytes = int(lib.cc_pcg_workspace_bytes(A.n, k, mode))
    ws = (workspace or _default_ws).get(nbytes, A.vals.device)
    res = (PcgResult * k)()
    hist = None
    if runs is not None and k != 1:
        raise ValueError("the run-encoded products serve one right-hand side")
    if runs is not None and mode == 0:
        if history:
            hist = torch.full((int(maxiter) + 1, k), float("nan"), dtype=torch.float64, device=A.vals.device)
        st = lib.cc_pcg_jacobi_f64_runs(
            A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(runs.rptr), _ptr(runs.runs), _ptr(B2), _ptr(X2),
            float(tol), int(maxiter), prec, int(check_every), _ptr(ws), nbytes, ctypes.addressof(res),
            _ptr(hist) if hist is not None else None, _stream(),
        )
    elif runs is not None:
        st = lib.cc_pcg_jacobi_mixed_runs(
            A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(A.ensure_f32()), _ptr(runs.rptr),
            _ptr(runs.runs), _ptr(B2), _ptr(X2), float(tol), int(maxiter), float(inner_tol), int(max_outer), prec,
            int(check_every), _ptr(ws), nbytes, ctypes.addressof(res), _stream(),
        )
    elif mode == 0:
        if history:
            hist = torch.full((int(maxiter) + 1, k), float("nan"), dtype=torch.float64, device=A.vals.device)
        if rb is not None:
            st = lib.cc_pcg_jacobi_f64_rb(
                A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), ctypes.addressof(rb.desc), _ptr(B2),
                _ptr(X2), float(tol), int(maxiter), prec, int(check_every), _ptr(ws), nbytes,
                ctypes.addressof(res), _ptr(hist) if hist is not None else None, _stream(),
            )
        else:
            st = lib.cc_pcg_jacobi_f64(
                A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(B2), _ptr(X2),
                float(tol), int(maxiter), prec, int(check_every), _ptr(ws), nbytes, ctypes.addressof(res),
                _ptr(hist) if hist is not None else None, _stream(),
            )
    elif rb is not None:
        st = lib.cc_pcg_jacobi_mixed_rb(
            A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), ctypes.addressof(rb.desc), _ptr(B2), _ptr(X2),
            float(tol), int(maxiter), float(inner_tol), int(max_outer), prec, int(check_every),
            _ptr(ws), nbytes, ctypes.addressof(res), _stream(),
        )
    else:
        v32 = A.ensure_f32()
        st = lib.cc_pcg_jacobi_mixed(
            A.n, k, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(v32), _ptr(B2), _ptr(X2),
            float(tol), int(maxiter), float(inner_tol), int(max_outer), prec, int(check_every),
            _ptr(ws), nbytes, ctypes.addressof(res), _stream(),
        )
    _check(st, "pcg")
    out = X2.reshape(-1) if squeeze else X2
    if history:
        return out, list(res), (hist.cpu().numpy() if hist is not None else None)
    return out, list(res)


def gmres(A: DeviceCSR, b, *, tol: float = 1.0e-8, restart: int = 80, maxiter: int = 5000,
          preconditioner: str = "jacobi", workspace: Workspace | None = None):
    """Restarted GMRES (float64, one right-hand side) on the device; returns
    ``(x, GmresResult)``. ``maxiter`` counts restarts, as SciPy's does."""
    lib = require_gpu()
    torch = _torch()
    bd = b.reshape(-1).to(torch.float64).contiguous()
    if bd.numel() != A.n:
        raise ValueError("b has the wrong length")
    x = torch.empty_like(bd)
    rs = int(min(restart if restart > 0 else 20, A.n))
    nbytes = int(lib.cc_gmres_workspace_bytes(A.n, rs))
    ws = (workspace or _default_ws).get(nbytes, A.vals.device)
    res = GmresResult()
    prec = {"none": 0, "jacobi": 1}[preconditioner]
    _check(lib.cc_gmres_jacobi_f64(A.n, _ptr(A.indptr), _ptr(A.indices), _ptr(A.vals), _ptr(bd), _ptr(x),
                                   float(tol), rs, int(maxiter), prec, _ptr(ws), nbytes, ctypes.addressof(res),
                                   _stream()), "gmres")
    return x, res
