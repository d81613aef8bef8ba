"""Write the reference-generated aggregation fixtures under ``tests/golden/``.

The extension rules pick each small dof's stencil by sorting Greville
distances and testing ranks at a tolerance, so the sparsity of ``E`` (and the
exact-zero cancellations that set the nnz of the restricted operator) depends
on tie order and last-bit rounding that no closed form or invariant pins.
This script therefore imports the reference's own functions (in this
container, from ``/root/reference/code``) and records their integer outputs
(active / interior / small index maps, the CSR pattern of ``E``, the nnz of
the aggregated operator) plus the float outputs needed to check values
(``gamma``, ``E`` entries, the operator). Only data is written; the
reference itself never travels with the repository.

Functions exercised (reference file:line):
- ``oblique_cut_sbp.assemble_oblique_cut_matrices`` (271-433),
  ``polynomial_preserving_oblique_extension`` (614-672),
  ``pullback_oblique_2d`` (693-707), with the sliver recipe of
  ``fast_solver_workbench.run_case`` (138-168);
- ``bell_3d_core.assemble_bell_3d`` (275-430),
  ``polynomial_preserving_extension_3d`` (471-542), ``restrict_spd``
  (550-552), with the system recipe of
  ``benchmarks/gpu_pcg_benchmark_3d.build_system`` (49-57).

Run: ``OMP_NUM_THREADS=1 python tests/golden/make_golden.py`` (about a
minute). Output: ``tests/golden/agg_<case>.npz`` and ``agg_manifest.json``.
"""

from __future__ import annotations

import json
import os
import platform
import sys
import time
from math import pi
from pathlib import Path

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import numpy as np  # noqa: E402
import scipy  # noqa: E402
import scipy.sparse as sp  # noqa: E402

REF = Path("/root/reference/code")
OUT = Path(__file__).resolve().parent
sys.path.insert(0, str(REF))

import bell_3d_core as b3  # noqa: E402
import oblique_cut_sbp as ob  # noqa: E402
from cut_spline_sbp import make_open_uniform_basis  # noqa: E402


def _csr_fields(prefix: str, M: sp.spmatrix, values: bool = True) -> dict:
    M = sp.csr_matrix(M)
    M.sort_indices()
    out = {
        f"{prefix}_shape": np.asarray(M.shape, dtype=np.int64),
        f"{prefix}_indptr": M.indptr.astype(np.int64),
        f"{prefix}_indices": M.indices.astype(np.int32),
    }
    if values:
        out[f"{prefix}_data"] = M.data.astype(np.float64)
    return out


def case_2d(name: str, n_elements: int, degree: int, angle: float, offset: float, active_tol: float) -> dict:
    """Dense 2D assembly + extension + ``A = csr(sym(E^T (K + H) E))``."""
    bx = make_open_uniform_basis(n_elements=n_elements, degree=degree)
    by = make_open_uniform_basis(n_elements=n_elements, degree=degree)
    cut = ob.HalfPlaneCut.from_angle(angle, offset)
    q = max(24, 2 * degree + 14)
    m = ob.assemble_oblique_cut_matrices(bx, by, cut, quadrature_order=q, capture_volume_quadrature=False)
    agg = ob.polynomial_preserving_oblique_extension(
        bx, by, m, support_threshold=0.05, preserve_degree=degree, active_tol=active_tol
    )
    phys = ob.pullback_oblique_2d(m, agg.E)
    A = sp.csr_matrix(0.5 * ((phys.K + phys.H) + (phys.K + phys.H).T))
    data = {
        "active": agg.active.astype(np.int64),
        "interior": agg.interior.astype(np.int64),
        "small": agg.small_support.astype(np.int64),
        "gamma": np.asarray(agg.gamma, dtype=np.float64),
        "polynomial_degree": np.asarray(agg.polynomial_degree),
        "polynomial_residual": np.asarray(agg.polynomial_residual),
        **_csr_fields("E", sp.csr_matrix(agg.E)),
        **_csr_fields("A", A),
    }
    meta = dict(kind="oblique2d", n_elements=n_elements, degree=degree, angle=angle, offset=offset,
                quadrature_order=q, support_threshold=0.05, active_tol=active_tol)
    return name, data, meta


def case_bell(name: str, n_elements: int, degree: int = 2, epsilon: float = 1.0e-3, store_A: bool = False) -> dict:
    geom = b3.Bell3DGeometry.from_grid(20, epsilon)
    asm = b3.assemble_bell_3d(geom, n_elements=n_elements, degree=degree, with_flux=False)
    agg = b3.polynomial_preserving_extension_3d(asm)
    A = (b3.restrict_spd(asm.K, agg.E) + b3.restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    data = {
        "active": np.asarray(agg.active, dtype=np.int64),
        "interior": np.asarray(agg.interior, dtype=np.int64),
        "small": np.asarray(agg.small_support, dtype=np.int64),
        "gamma": np.asarray(agg.gamma, dtype=np.float64),
        "polynomial_residual": np.asarray(agg.polynomial_residual),
        **_csr_fields("E", agg.E),
        "A_nnz": np.asarray(A.nnz, dtype=np.int64),
        "A_shape": np.asarray(A.shape, dtype=np.int64),
        "A_row_nnz": np.diff(A.indptr).astype(np.int32),
    }
    if store_A:
        data.update(_csr_fields("A", A, values=False))
    meta = dict(kind="bell3d", n_elements=n_elements, degree=degree, epsilon=epsilon, geometry_n=20,
                support_threshold=0.04)
    return name, data, meta


def main() -> None:
    rng = np.random.default_rng(0)
    angles = rng.uniform(0.0, 45.0, 72)
    offsets = rng.uniform(0.05, 0.95, 72)
    jobs = []
    # Sliver stress recipe (vertical cut at x = h (1 - eps), active_tol 1e-100).
    for N, eps in ((8, 1.0e-3), (16, 1.0e-5), (32, 1.0e-1)):
        jobs.append((case_2d, (f"sliver_N{N}_p4_eps{eps:.0e}", N, 4, 0.0, (1.0 / N) * (1.0 - eps), 1.0e-100)))
    # The small 2D configuration: 5 degree cut at offset 0.85 h, N = 32, p = 4.
    jobs.append((case_2d, ("oblique_N32_p4_5deg", 32, 4, 5.0 * pi / 180.0, 0.85 / 32, 1.0e-22)))
    # The first random cut of the batched 2D configuration (N = 48), each degree.
    for p in (3, 4, 5):
        jobs.append((case_2d, (f"oblique_N48_p{p}_cut0", 48, p, float(angles[0]) * pi / 180.0,
                               float(offsets[0]) / 48, 1.0e-22)))
    # The reference's published 3D bell family (p = 2, eps = 1e-3).
    jobs.append((case_bell, ("bell_N12_p2", 12, 2, 1.0e-3, True)))
    jobs.append((case_bell, ("bell_N20_p2", 20)))

    manifest = {
        "generator": "tests/golden/make_golden.py",
        "numpy": np.__version__,
        "scipy": scipy.__version__,
        "python": platform.python_version(),
        "machine": platform.machine(),
        "cases": {},
    }
    for fn, args in jobs:
        t0 = time.perf_counter()
        name, data, meta = fn(*args)
        path = OUT / f"agg_{name}.npz"
        np.savez_compressed(path, **data)
        meta.update(n_active=int(data["active"].size), n_interior=int(data["interior"].size),
                    n_small=int(data["small"].size), E_nnz=int(data["E_indices"].size),
                    A_nnz=int(data["A_nnz"]) if "A_nnz" in data else int(data["A_indices"].size),
                    bytes=path.stat().st_size)
        manifest["cases"][name] = meta
        print(f"{name}: {time.perf_counter() - t0:.1f} s {meta}", flush=True)
    (OUT / "agg_manifest.json").write_text(json.dumps(manifest, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
