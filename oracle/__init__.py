"""CPU oracle of the cut-cell Jacobi-PCG path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for the
CPU baseline, as the timed CPU port). The product path (``fast_solver_backends``
and ``cutcell``) never imports it.

Contents:
- ``pcg_ref.c`` (built into ``liboracle.so`` by ``Makefile``): scalar float64
  restatements of the reference GPU loop (``code/fast_solver_backends.py:233-274``)
  and of SciPy 1.15.3 CG as called by ``solve_cpu_cg``
  (``code/fast_solver_backends.py:90-132``; SciPy ``iterative.py:375-423``).
- ``pcg_gpu_loop_py``: the same GPU-loop recurrence in NumPy, a second,
  independent restatement used to pin the C one.
- ``mixed_refinement_py``: float32 inner Jacobi-PCG under float64 iterative
  refinement, the recurrence the HIP mixed-precision solve follows.

How each function is pinned (see the tests):
- ``oracle_pcg_scipy``: iteration counts equal SciPy's own ``cg`` and the
  reference's committed ``code/results/fast_solver_stress.csv`` on its sliver
  systems; closed-form one-step convergence on diagonal systems.
- ``oracle_pcg_gpu_loop``: within one iteration of ``oracle_pcg_scipy`` on every
  stress system (the reference's own GPU-vs-CPU agreement rule,
  ``code/tests/test_headline_results.py:90-97``); hand-computed first iterates
  on 2x2 systems; breakdown on ``A = -I`` and ``b = 0``; ``maxiter``
  exhaustion; agreement with the NumPy restatement; solutions against
  SuperLU.
- ``mixed_refinement_py``: true float64 residual below ``tol`` and solutions
  against SuperLU; with ``inner_tol`` tiny and exact float64 data it reduces to
  one sweep.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

CONVERGED, BREAKDOWN, MAXITER = 1, 2, 3

_lib = None


def build() -> Path:
    """Compile ``pcg_ref.c`` with gcc (no GPU, no reference sources needed)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        P, I, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        lib.oracle_pcg_gpu_loop.argtypes = [I, P, P, P, P, P, D, I, I, P, P, P, P]
        lib.oracle_pcg_gpu_loop.restype = ctypes.c_int
        lib.oracle_pcg_scipy.argtypes = [I, P, P, P, P, P, D, I, I, P, P]
        lib.oracle_pcg_scipy.restype = ctypes.c_int
        lib.oracle_spmv.argtypes = [I, P, P, P, P, P]
        lib.oracle_spmv.restype = None
        _lib = lib
    return _lib


def _csr_arrays(A):
    A = A.tocsr()
    return (
        np.ascontiguousarray(A.indptr, dtype=np.int32),
        np.ascontiguousarray(A.indices, dtype=np.int32),
        np.ascontiguousarray(A.data, dtype=np.float64),
    )


def spmv(A, x: np.ndarray) -> np.ndarray:
    lib = _load()
    ip, ix, v = _csr_arrays(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(A.shape[0])
    lib.oracle_spmv(A.shape[0], ip.ctypes.data, ix.ctypes.data, v.ctypes.data, x.ctypes.data, y.ctypes.data)
    return y


def pcg_gpu_loop(A, b, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi", history=False):
    """Reference GPU-loop Jacobi-PCG. Returns ``(x, iterations, status, rel)``
    (plus the residual history ``||r_k||/||b||``, k = 0..iterations, when asked)."""
    lib = _load()
    ip, ix, v = _csr_arrays(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int32()
    st = ctypes.c_int32()
    rel = ctypes.c_double()
    hist = np.full(int(maxiter) + 1, np.nan) if history else None
    rc = lib.oracle_pcg_gpu_loop(
        A.shape[0], ip.ctypes.data, ix.ctypes.data, v.ctypes.data, b.ctypes.data, x.ctypes.data,
        float(tol), int(maxiter), int(preconditioner == "jacobi"),
        ctypes.addressof(it), ctypes.addressof(st), ctypes.addressof(rel),
        hist.ctypes.data if history else None,
    )
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    if history:
        return x, int(it.value), int(st.value), float(rel.value), hist[: int(it.value) + 1]
    return x, int(it.value), int(st.value), float(rel.value)


def pcg_scipy(A, b, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi"):
    """SciPy-CG semantics. Returns ``(x, iterations, status)``."""
    lib = _load()
    ip, ix, v = _csr_arrays(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    it = ctypes.c_int32()
    st = ctypes.c_int32()
    rc = lib.oracle_pcg_scipy(
        A.shape[0], ip.ctypes.data, ix.ctypes.data, v.ctypes.data, b.ctypes.data, x.ctypes.data,
        float(tol), int(maxiter), int(preconditioner == "jacobi"), ctypes.addressof(it), ctypes.addressof(st),
    )
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return x, int(it.value), int(st.value)


def pcg_gpu_loop_py(A, b, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi"):
    """NumPy restatement of the reference GPU loop (small systems)."""
    A = A.tocsr()
    b = np.asarray(b, dtype=float)
    d = A.diagonal()
    jac = preconditioner == "jacobi"

    def prec(r):
        if not jac:
            return r.copy()
        z = r.copy()
        m = np.abs(d) > 0.0
        z[m] = r[m] / d[m]
        return z

    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rho = float(r @ z)
    bn = float(np.linalg.norm(b)) or 1.0
    status, iters = MAXITER, 0
    rel = float(np.linalg.norm(r)) / bn
    for k in range(1, maxiter + 1):
        q = A @ p
        den = float(p @ q)
        if den <= 0.0 or not np.isfinite(den):
            status, iters = BREAKDOWN, k
            break
        alpha = rho / den
        x += alpha * p
        r -= alpha * q
        rel = float(np.linalg.norm(r)) / bn
        iters = k
        if rel <= tol:
            status = CONVERGED
            break
        z = prec(r)
        rho_new = float(r @ z)
        p = z + (rho_new / rho) * p
        rho = rho_new
    return x, iters, status, rel


def mixed_refinement_py(A, b, *, tol=1.0e-8, maxiter=5000, inner_tol=1.0e-4, max_outer=12):
    """Float32 inner Jacobi-PCG (float64 dots) under float64 refinement.

    Returns ``(x, total_inner_iterations, sweeps, status, rel)``.
    """
    A = A.tocsr()
    A32 = A.astype(np.float32)
    d = A.diagonal()
    b = np.asarray(b, dtype=float)
    bn = float(np.linalg.norm(b)) or 1.0
    x = np.zeros_like(b)
    r = b - A @ x
    total = 0
    sweeps = 0
    rel = float(np.linalg.norm(r)) / bn
    while rel > tol and sweeps < max_outer and total < maxiter:
        scale = float(np.linalg.norm(r))
        rr32 = (r / scale).astype(np.float32)
        xi = np.zeros_like(rr32)
        ri = rr32.copy()

        def prec(v):
            out = v.copy()
            m = np.abs(d) > 0.0
            out[m] = (v[m].astype(np.float64) / d[m]).astype(np.float32)
            return out

        zi = prec(ri)
        pi = zi.copy()
        rho = float(np.dot(ri.astype(np.float64), zi.astype(np.float64)))
        bni = float(np.linalg.norm(ri.astype(np.float64))) or 1.0
        its = 0
        for k in range(1, maxiter - total + 1):
            qi = (A32 @ pi).astype(np.float32)
            den = float(np.dot(pi.astype(np.float64), qi.astype(np.float64)))
            if den <= 0.0 or not np.isfinite(den):
                its = k
                break
            a = rho / den
            xi = (xi.astype(np.float64) + a * pi.astype(np.float64)).astype(np.float32)
            ri = (ri.astype(np.float64) - a * qi.astype(np.float64)).astype(np.float32)
            its = k
            if np.linalg.norm(ri.astype(np.float64)) / bni <= inner_tol:
                break
            zi = prec(ri)
            rn = float(np.dot(ri.astype(np.float64), zi.astype(np.float64)))
            pi = (zi.astype(np.float64) + (rn / rho) * pi.astype(np.float64)).astype(np.float32)
            rho = rn
        total += its
        sweeps += 1
        x = x + scale * xi.astype(np.float64)
        r = b - A @ x
        rel = float(np.linalg.norm(r)) / bn
    status = CONVERGED if rel <= tol else MAXITER
    return x, total, sweeps, status, rel
