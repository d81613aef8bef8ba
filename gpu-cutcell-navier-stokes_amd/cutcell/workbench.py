"""The sliver stress workbench around ``solve_spd``: cases, result rows and
their files, diffable against the reference's own results.

Reference behaviour kept (reference repository, file:line):
- ``SolverCase`` / ``SolverRow``: ``code/fast_solver_workbench.py:72-107``
  (same fields, names and order);
- ``run_case``: the sliver system (vertical cut at ``x = h(1 - eps)``,
  ``A = sym(E^T K E + E^T H E)``, Gaussian right-hand side with seed
  ``12345 + 17 N + p``), one ``solve_spd`` call with the Jacobi
  preconditioner, the optional condition proxy (dense eigenvalues up to 900
  unknowns, ``eigsh`` at both ends above): ``:114-198``;
- one case per worker process, rows sorted by (degree, N, eps): ``:201-217``;
- ``write_outputs``: ``{prefix}.json`` (``{"rows", "summary"}``),
  ``{prefix}_summary.json`` and ``{prefix}.csv`` with the same summary keys
  and column order: ``:220-244`` (the reference's figure is not written).

The system builder is this package's (``cutcell.oblique2d.sliver_system``);
its integer sizes equal the reference's committed rows exactly.
"""

from __future__ import annotations

import csv
import json
import os
from concurrent.futures import ProcessPoolExecutor
from dataclasses import asdict, dataclass, fields
from pathlib import Path
from time import perf_counter

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


@dataclass(frozen=True)
class SolverCase:
    degree: int
    n_elements: int
    epsilon: float
    backend: str
    tol: float
    maxiter: int
    compute_condition: bool


@dataclass(frozen=True)
class SolverRow:
    degree: int
    n_elements: int
    epsilon: float
    volume_ratio: float
    n_active: int
    n_reduced: int
    nnz: int
    csr_memory_mb: float
    assembly_s: float
    solve_s: float
    transfer_s: float
    iterations: int
    success: bool
    relative_residual: float
    condition_proxy: float | None
    backend: str
    method: str
    preconditioner: str
    message: str


def condition_proxy(A: sp.csr_matrix) -> float:
    """Spectral condition number (dense up to 900 unknowns, else the two
    ends by ``eigsh``); inf when the smallest eigenvalue is not positive."""
    n = A.shape[0]
    if n <= 900:
        ev = np.linalg.eigvalsh(A.toarray())
        return float(ev[-1] / ev[0]) if ev[0] > 0.0 else float("inf")
    try:
        hi = float(spla.eigsh(A, k=1, which="LA", return_eigenvectors=False, tol=1.0e-4)[0])
        lo = float(spla.eigsh(A, k=1, which="SA", return_eigenvectors=False, tol=1.0e-4)[0])
        return float(hi / lo) if lo > 0.0 else float("inf")
    except Exception:
        return float("nan")


def run_case(case: SolverCase) -> SolverRow:
    from fast_solver_backends import csr_memory_bytes, solve_spd

    from .oblique2d import sliver_system

    t0 = perf_counter()
    s = sliver_system(case.n_elements, case.degree, case.epsilon)
    assembly_s = perf_counter() - t0
    _, rep = solve_spd(s.A, s.b, backend=case.backend, tol=case.tol, maxiter=case.maxiter, preconditioner="jacobi")
    cond = condition_proxy(s.A) if case.compute_condition else None
    return SolverRow(
        degree=int(case.degree),
        n_elements=int(case.n_elements),
        epsilon=float(case.epsilon),
        volume_ratio=float(1.0 / case.epsilon),
        n_active=int(len(s.active)),
        n_reduced=int(s.E.shape[1]),
        nnz=int(s.A.nnz),
        csr_memory_mb=float(csr_memory_bytes(s.A) / 1.0e6),
        assembly_s=float(assembly_s),
        solve_s=float(rep.elapsed_s),
        transfer_s=float(rep.transfer_s),
        iterations=int(rep.iterations),
        success=bool(rep.success),
        relative_residual=float(rep.relative_residual),
        condition_proxy=None if cond is None else float(cond),
        backend=rep.backend,
        method=rep.method,
        preconditioner=rep.preconditioner,
        message=rep.message,
    )


def run_cases(cases, jobs: int | None = None) -> list[SolverRow]:
    """Every case, sorted as the reference sorts its rows. GPU cases run in
    this process (one GPU, one solve at a time); CPU cases over ``jobs``
    worker processes with one BLAS thread each."""
    cases = list(cases)
    gpu = [c for c in cases if c.backend != "cpu"]
    cpu = [c for c in cases if c.backend == "cpu"]
    rows = [run_case(c) for c in gpu]
    jobs = jobs or min(16, os.cpu_count() or 1)
    if cpu:
        if jobs <= 1 or len(cpu) == 1:
            rows += [run_case(c) for c in cpu]
        else:
            for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
                os.environ.setdefault(var, "1")
            with ProcessPoolExecutor(max_workers=jobs) as ex:
                rows += list(ex.map(run_case, cpu))
    rows.sort(key=lambda r: (r.degree, r.n_elements, r.epsilon))
    return rows


def summarize(rows: list[SolverRow]) -> dict:
    """The reference's summary record (same keys and order)."""
    return {
        "n_cases": len(rows),
        "backend_values": sorted({r.backend for r in rows}),
        "degree_values": sorted({r.degree for r in rows}),
        "n_elements_values": sorted({r.n_elements for r in rows}),
        "volume_ratio_range": [min((r.volume_ratio for r in rows), default=None),
                               max((r.volume_ratio for r in rows), default=None)],
        "all_converged": all(r.success for r in rows),
        "iteration_range": [min((r.iterations for r in rows), default=None),
                            max((r.iterations for r in rows), default=None)],
        "solve_time_s_range": [min((r.solve_s for r in rows), default=None), max((r.solve_s for r in rows), default=None)],
        "assembly_time_s_range": [min((r.assembly_s for r in rows), default=None),
                                  max((r.assembly_s for r in rows), default=None)],
        "max_relative_residual": max((r.relative_residual for r in rows), default=None),
        "max_csr_memory_mb": max((r.csr_memory_mb for r in rows), default=None),
    }


def write_outputs(rows: list[SolverRow], prefix: str, directory: Path | str) -> dict:
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    data = [asdict(r) for r in rows]
    summary = summarize(rows)
    (directory / f"{prefix}.json").write_text(json.dumps({"rows": data, "summary": summary}, indent=2),
                                              encoding="utf-8")
    (directory / f"{prefix}_summary.json").write_text(json.dumps(summary, indent=2), encoding="utf-8")
    with (directory / f"{prefix}.csv").open("w", newline="", encoding="utf-8") as f:
        writer = csv.DictWriter(f, fieldnames=[fl.name for fl in fields(SolverRow)])
        writer.writeheader()
        writer.writerows(data)
    return summary


def _coerce(name: str, value):
    kind = {fl.name: fl.type for fl in fields(SolverRow)}[name]
    if name == "condition_proxy":
        return None if value in (None, "") else float(value)
    if kind == "int":
        return int(value)
    if kind == "float":
        return float(value)
    if kind == "bool":
        return value if isinstance(value, bool) else str(value) == "True"
    return str(value)


def row_from_dict(d: dict) -> SolverRow:
    return SolverRow(**{fl.name: _coerce(fl.name, d[fl.name]) for fl in fields(SolverRow)})


def read_rows(path: Path | str) -> list[SolverRow]:
    """Rows of a results file written by the reference or by ``write_outputs``
    (``.csv``, or ``.json`` with a ``rows`` list)."""
    path = Path(path)
    if path.suffix == ".json":
        return [row_from_dict(d) for d in json.loads(path.read_text(encoding="utf-8"))["rows"]]
    with path.open(newline="", encoding="utf-8") as f:
        return [row_from_dict(d) for d in csv.DictReader(f)]


def stress_cases(backend: str = "cpu", tol: float = 1.0e-8, maxiter: int = 5000, compute_condition: bool = False):
    """The reference's committed sweep: p = 4, N in {8, 12, 16, 24, 32},
    eps in {1e-5, ..., 1e-1}."""
    return [SolverCase(4, N, eps, backend, tol, maxiter, compute_condition)
            for N in (8, 12, 16, 24, 32) for eps in (1.0e-5, 1.0e-4, 1.0e-3, 1.0e-2, 1.0e-1)]
