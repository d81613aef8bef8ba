"""Drop-in solver backends for aggregated cut-cell SPD systems.

Same public surface as the reference module ``code/fast_solver_backends.py``
(``SolverReport`` 30-45, ``as_csr``/``csr_memory_bytes`` 48-60,
``relative_residual`` 63-71, ``jacobi_preconditioner`` 74-87, ``solve_cpu_cg``
90-132, ``solve_cpu_gmres`` 135-191, ``solve_gpu_cg`` 194-289, ``solve_spd``
292-325), so callers switch by changing the import.

``backend="gpu"`` runs on an AMD Instinct GPU through the hand-written gfx950
HIP kernels of ``cutcell/libcutcell_hip.so`` (method ``"hip-jacobi-pcg"``):
the CSR matrix and right-hand side are copied to HBM once, the whole CG loop
runs on the device with no per-iteration host synchronisation, and the
report's ``relative_residual`` is the float64 true residual recomputed on the
device. The CPU backend calls SciPy exactly as the reference does.

Behaviour kept from the reference:
- ``solve_gpu_cg`` returns ``(None, report(success=False))`` when no GPU is
  visible ("HIP backend unavailable: ..."); ``solve_spd(backend="auto")``
  then reruns on the CPU with the message prefix ``"GPU skipped/fell back: "``.
  A process that sees a GPU but cannot load the HIP library raises instead.
- GPU success requires convergence of the recurrence and a true residual
  ``<= max(10 tol, 1e-14)``; the iteration count, stopping test and breakdown
  test are those of the reference GPU loop.
- ``elapsed_s`` excludes the host-to-device copy (``transfer_s``) and includes
  the final device synchronisation.
"""

from __future__ import annotations

from dataclasses import asdict, dataclass
from time import perf_counter
from typing import Any, Literal

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

Backend = Literal["cpu", "gpu"]
Preconditioner = Literal["none", "jacobi"]
Precision = Literal["fp64", "mixed"]


@dataclass(frozen=True)
class SolverReport:
    """Summary returned by every solver backend."""

    backend: str
    method: str
    preconditioner: str
    success: bool
    iterations: int
    elapsed_s: float
    relative_residual: float
    message: str = ""
    transfer_s: float = 0.0

    def to_dict(self) -> dict[str, Any]:
        return asdict(self)


def as_csr(A) -> sp.csr_matrix:
    """``A`` as CSR without changing its values (dense input becomes float)."""
    if sp.issparse(A):
        return A.tocsr()
    return sp.csr_matrix(np.asarray(A, dtype=float))


def csr_memory_bytes(A) -> int:
    """CSR storage in bytes: values + column indices + row pointers."""
    Acsr = as_csr(A)
    return int(Acsr.data.nbytes + Acsr.indices.nbytes + Acsr.indptr.nbytes)


def relative_residual(A, x: np.ndarray, b: np.ndarray) -> float:
    """``||Ax - b|| / ||b||`` with ``||b|| = 0 -> 1``."""
    Acsr = as_csr(A)
    b = np.asarray(b, dtype=float)
    denom = float(np.linalg.norm(b))
    if denom == 0.0:
        denom = 1.0
    return float(np.linalg.norm(Acsr @ x - b) / denom)


def jacobi_preconditioner(A) -> spla.LinearOperator:
    """``z = D^{-1} r`` with zero diagonal entries mapped to 0 (CPU rule)."""
    Acsr = as_csr(A)
    d = Acsr.diagonal().astype(float, copy=True)
    inv = np.zeros_like(d)
    mask = np.abs(d) > 0.0
    inv[mask] = 1.0 / d[mask]
    return spla.LinearOperator(Acsr.shape, matvec=lambda x: inv * x, dtype=float)


def _counted(fn, *args, **kwargs):
    count = 0

    def callback(*_):
        nonlocal count
        count += 1

    out = fn(*args, callback=callback, **kwargs)
    return out, count


def solve_cpu_cg(A, b, *, tol: float = 1.0e-8, maxiter: int = 5000, preconditioner: Preconditioner = "jacobi"):
    """SciPy CG (``rtol=tol, atol=0``); the iteration count is the callback count."""
    Acsr = as_csr(A)
    b = np.asarray(b, dtype=float)
    M = jacobi_preconditioner(Acsr) if preconditioner == "jacobi" else None
    tic = perf_counter()
    (x, info), count = _counted(spla.cg, Acsr, b, rtol=tol, atol=0.0, maxiter=maxiter, M=M)
    elapsed = perf_counter() - tic
    rel = relative_residual(Acsr, x, b)
    success = bool(info == 0 and np.isfinite(rel) and rel <= max(10.0 * tol, 1.0e-14))
    if info == 0:
        message = "converged"
    elif info > 0:
        message = f"not converged after {info} iterations"
    else:
        message = f"illegal input or breakdown (info={info})"
    return x, SolverReport(
        backend="cpu",
        method="scipy.sparse.linalg.cg",
        preconditioner=preconditioner,
        success=success,
        iterations=int(count if count else max(info, 0)),
        elapsed_s=float(elapsed),
        relative_residual=float(rel),
        message=message,
    )


def solve_cpu_gmres(
    A, b, *, tol: float = 1.0e-8, maxiter: int = 5000, restart: int = 80, preconditioner: Preconditioner = "jacobi"
):
    """Restarted SciPy GMRES, counted per inner iteration (``pr_norm`` callback)."""
    Acsr = as_csr(A)
    b = np.asarray(b, dtype=float)
    M = jacobi_preconditioner(Acsr) if preconditioner == "jacobi" else None
    tic = perf_counter()
    (x, info), count = _counted(
        spla.gmres, Acsr, b, rtol=tol, atol=0.0, restart=restart, maxiter=maxiter, M=M, callback_type="pr_norm"
    )
    elapsed = perf_counter() - tic
    rel = relative_residual(Acsr, x, b)
    success = bool(info == 0 and np.isfinite(rel) and rel <= max(10.0 * tol, 1.0e-14))
    message = "converged" if info == 0 else f"not converged or breakdown (info={info})"
    return x, SolverReport(
        backend="cpu",
        method="scipy.sparse.linalg.gmres",
        preconditioner=preconditioner,
        success=success,
        iterations=int(count),
        elapsed_s=float(elapsed),
        relative_residual=float(rel),
        message=message,
    )


def _unavailable(preconditioner: str, why: str, method: str = "hip-jacobi-pcg"):
    return None, SolverReport(
        backend="gpu",
        method=method,
        preconditioner=preconditioner,
        success=False,
        iterations=0,
        elapsed_s=0.0,
        relative_residual=float("inf"),
        message=f"HIP backend unavailable: {why}",
    )


def solve_gpu_cg(
    A,
    b,
    *,
    tol: float = 1.0e-8,
    maxiter: int = 5000,
    preconditioner: Preconditioner = "jacobi",
    precision: Precision = "fp64",
    check_every: int = 16,
):
    """Jacobi-PCG on the GPU through the gfx950 HIP kernels.

    ``precision="mixed"`` keeps float32 values and vectors in an inner PCG and
    recomputes the residual, the dot products and the correction in float64
    (iterative refinement); the stopping test is then on the float64 true
    residual and ``iterations`` counts inner CG iterations.
    """
    method = "hip-jacobi-pcg" if precision == "fp64" else "hip-mixed-pcg"
    try:
        import torch
    except Exception as exc:  # pragma: no cover - torch is part of the image
        return _unavailable(preconditioner, repr(exc), method)
    if not torch.cuda.is_available():
        return _unavailable(preconditioner, "no ROCm GPU visible", method)
    from cutcell import hip  # raises loudly when a GPU is present but the library is not

    hip.load_library()
    Acsr = as_csr(A)
    b_np = np.asarray(b, dtype=float)
    t0 = perf_counter()
    Ad = hip.DeviceCSR.from_scipy(Acsr, with_f32=(precision == "mixed"))
    bd = torch.from_numpy(np.ascontiguousarray(b_np)).to("cuda")
    torch.cuda.synchronize()
    transfer_s = perf_counter() - t0

    t0 = perf_counter()
    # The products come from the run encoding of the rows (built on the
    # device, counted in elapsed_s) when the column indices ascend in each
    # row, as SciPy's canonical CSR has them; otherwise from the CSR arrays.
    runs = hip.DeviceRuns.from_csr(Ad) if Acsr.has_sorted_indices and Acsr.has_canonical_format else None
    xd, res = hip.pcg(
        Ad, bd, tol=tol, maxiter=maxiter, preconditioner=preconditioner, precision=precision, check_every=check_every,
        runs=runs,
    )
    torch.cuda.synchronize()
    elapsed = perf_counter() - t0
    r = res[0]
    x_np = xd.cpu().numpy()
    rel = float(r.true_rel)
    converged = r.status == hip.CC_CONVERGED
    return x_np, SolverReport(
        backend="gpu",
        method=method,
        preconditioner=preconditioner,
        success=bool(converged and np.isfinite(rel) and rel <= max(10.0 * tol, 1.0e-14)),
        iterations=int(r.iterations),
        elapsed_s=float(elapsed),
        relative_residual=rel,
        message=hip.STATUS_MESSAGE[int(r.status)],
        transfer_s=float(transfer_s),
    )


def solve_gpu_gmres(
    A, b, *, tol: float = 1.0e-8, maxiter: int = 5000, restart: int = 80, preconditioner: Preconditioner = "jacobi"
):
    """Restarted GMRES on the GPU through the gfx950 HIP kernels: the
    semantics of ``solve_cpu_gmres`` (SciPy's gmres, the Jacobi CPU rule,
    iterations counted per inner step, the same success rule), for
    nonsymmetric operators such as convection-diffusion."""
    method = "hip-gmres"
    try:
        import torch
    except Exception as exc:  # pragma: no cover - torch is part of the image
        return _unavailable(preconditioner, repr(exc), method)
    if not torch.cuda.is_available():
        return _unavailable(preconditioner, "no ROCm GPU visible", method)
    from cutcell import hip

    hip.load_library()
    Acsr = as_csr(A)
    b_np = np.asarray(b, dtype=float)
    t0 = perf_counter()
    Ad = hip.DeviceCSR.from_scipy(Acsr)
    bd = torch.from_numpy(np.ascontiguousarray(b_np)).to("cuda")
    torch.cuda.synchronize()
    transfer_s = perf_counter() - t0
    t0 = perf_counter()
    xd, res = hip.gmres(Ad, bd, tol=tol, restart=restart, maxiter=maxiter, preconditioner=preconditioner)
    torch.cuda.synchronize()
    elapsed = perf_counter() - t0
    x_np = xd.cpu().numpy()
    rel = relative_residual(Acsr, x_np, b_np)
    info = int(res.info)
    return x_np, SolverReport(
        backend="gpu",
        method=method,
        preconditioner=preconditioner,
        success=bool(info == 0 and np.isfinite(rel) and rel <= max(10.0 * tol, 1.0e-14)),
        iterations=int(res.iterations),
        elapsed_s=float(elapsed),
        relative_residual=float(rel),
        message="converged" if info == 0 else f"not converged or breakdown (info={info})",
        transfer_s=float(transfer_s),
    )


def solve_spd(
    A,
    b,
    *,
    backend: Literal["cpu", "gpu", "auto"] = "cpu",
    tol: float = 1.0e-8,
    maxiter: int = 5000,
    preconditioner: Preconditioner = "jacobi",
    **gpu_options,
):
    """Dispatch to the requested SPD solver backend (``auto``: GPU, else CPU)."""
    if backend == "gpu":
        return solve_gpu_cg(A, b, tol=tol, maxiter=maxiter, preconditioner=preconditioner, **gpu_options)
    if backend == "auto":
        x, report = solve_gpu_cg(A, b, tol=tol, maxiter=maxiter, preconditioner=preconditioner, **gpu_options)
        if report.success:
            return x, report
        x_cpu, rc = solve_cpu_cg(A, b, tol=tol, maxiter=maxiter, preconditioner=preconditioner)
        message = rc.message
        if report.message:
            message = f"GPU skipped/fell back: {report.message}; CPU: {message}"
        return x_cpu, SolverReport(
            backend=rc.backend,
            method=rc.method,
            preconditioner=rc.preconditioner,
            success=rc.success,
            iterations=rc.iterations,
            elapsed_s=rc.elapsed_s,
            relative_residual=rc.relative_residual,
            message=message,
            transfer_s=rc.transfer_s,
        )
    return solve_cpu_cg(A, b, tol=tol, maxiter=maxiter, preconditioner=preconditioner)
