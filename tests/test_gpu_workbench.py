"""The reference's stress sweep (``code/fast_solver_workbench.py``) run
through ``cutcell.workbench`` on the HIP backend: every row's integer sizes
equal the reference's committed CSV, every solve succeeds with iterations
within ``parity.REFERENCE_ITER_GAP`` of the reference's, and the written
files read back as the same rows."""

from pathlib import Path

import pytest

from cutcell.workbench import read_rows, run_cases, stress_cases, write_outputs
from parity import REFERENCE_ITER_GAP

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REF_CSV = Path(__file__).resolve().parent / "golden" / "reference_fast_solver_stress.csv"


def test_stress_sweep_on_the_gpu_backend(tmp_path):
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    ref = {(r.n_elements, r.epsilon): r for r in read_rows(REF_CSV)}
    rows = run_cases(stress_cases(backend="gpu"))
    assert len(rows) == 25
    for r in rows:
        q = ref[(r.n_elements, r.epsilon)]
        print(f"N={r.n_elements} eps={r.epsilon:g}: {r.iterations} it (reference {q.iterations}), "
              f"rel {r.relative_residual:.2e}, {r.method}", flush=True)
        assert (r.n_active, r.n_reduced, r.nnz) == (q.n_active, q.n_reduced, q.nnz)
        assert r.backend == "gpu" and r.method == "hip-jacobi-pcg" and r.success and r.message == "converged"
        assert abs(r.iterations - q.iterations) <= REFERENCE_ITER_GAP
        assert r.relative_residual <= 1e-7 and r.transfer_s > 0
    summary = write_outputs(rows, "stress_gpu", tmp_path)
    assert summary["all_converged"] and summary["backend_values"] == ["gpu"]
    assert read_rows(tmp_path / "stress_gpu.csv") == rows
