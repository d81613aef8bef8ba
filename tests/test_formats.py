"""The workbench rows and files (``cutcell.workbench``) against the
reference's committed results, and the versioned system format
(``cutcell.formats``) on round trips and on the committed fixtures."""

import csv
import json
from pathlib import Path

import numpy as np
import pytest

from cutcell import formats
from cutcell.oblique2d import sliver_system
from cutcell.workbench import SolverCase, read_rows, run_cases, summarize, write_outputs
from parity import REFERENCE_ITER_GAP

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_CSV = GOLDEN / "reference_fast_solver_stress.csv"


def test_reference_rows_read_write_round_trip(tmp_path):
    rows = read_rows(REF_CSV)
    assert len(rows) == 25 and rows[0].method == "scipy.sparse.linalg.cg" and rows[0].condition_proxy > 0
    summary = write_outputs(rows, "stress", tmp_path)
    # The reference's own summary of these rows, value for value.
    ref_summary = json.loads((GOLDEN / "reference_fast_solver_stress_summary.json").read_text())
    assert list(summary) == list(ref_summary)
    assert summary == ref_summary
    header = REF_CSV.read_text().splitlines()[0]
    assert (tmp_path / "stress.csv").read_text().splitlines()[0] == header
    assert read_rows(tmp_path / "stress.csv") == rows
    assert read_rows(tmp_path / "stress.json") == rows
    assert json.loads((tmp_path / "stress_summary.json").read_text()) == ref_summary


def test_workbench_cpu_rows_match_the_reference():
    """A sample of the sweep through run_case on the CPU backend: integer
    sizes equal the reference's rows, iterations within the rounding gap."""
    ref = {(r.n_elements, r.epsilon): r for r in read_rows(REF_CSV)}
    cases = [SolverCase(4, N, eps, "cpu", 1e-8, 5000, False) for N in (8, 16) for eps in (1e-5, 1e-1)]
    rows = run_cases(cases, jobs=2)
    assert [(r.n_elements, r.epsilon) for r in rows] == sorted((c.n_elements, c.epsilon) for c in cases)
    for r in rows:
        q = ref[(r.n_elements, r.epsilon)]
        assert (r.n_active, r.n_reduced, r.nnz, r.csr_memory_mb) == (q.n_active, q.n_reduced, q.nnz, q.csr_memory_mb)
        assert r.success and r.message == "converged" and r.backend == "cpu"
        assert abs(r.iterations - q.iterations) <= REFERENCE_ITER_GAP
    s = summarize(rows)
    assert s["n_cases"] == 4 and s["all_converged"]


def test_system_round_trip_is_bit_exact(tmp_path):
    s = sliver_system(8, 4, 1e-3)
    path = formats.save_cut_system(tmp_path / "sliver.npz", s)
    got = formats.load_system(path)
    assert got.version == formats.VERSION and got.meta["kind"] == "sliver"
    d = formats.diff_systems(formats.from_cut_system(s), got)
    assert d["active_equal"] and d["interior_equal"] and d["small_equal"]
    assert d["E_pattern_equal"] and d["A_pattern_equal"]
    assert d["E_max_rel_diff"] == 0.0 and d["A_max_rel_diff"] == 0.0
    np.testing.assert_array_equal(got.b, s.b)


@pytest.mark.parametrize("name,N,eps", [("agg_sliver_N8_p4_eps1e-03.npz", 8, 1e-3),
                                        ("agg_sliver_N16_p4_eps1e-05.npz", 16, 1e-5)])
def test_reference_fixture_reads_as_version_0(name, N, eps):
    """The committed reference-generated fixtures read as version 0 and diff
    clean against the systems built here (index maps, E and A sparsity)."""
    ref = formats.load_system(GOLDEN / name)
    assert ref.version == 0 and ref.A_pattern is not None
    d = formats.diff_systems(ref, formats.from_cut_system(sliver_system(N, 4, eps)))
    assert d["active_equal"] and d["interior_equal"] and d["small_equal"]
    assert d["E_pattern_equal"] and d["A_pattern_equal"] and d["E_max_rel_diff"] < 1e-9
    assert d["A_max_rel_diff"] < 1e-11  # the 2D fixtures also hold the operator's values
    bell = formats.load_system(GOLDEN / "agg_bell_N12_p2.npz")
    assert bell.version == 0 and bell.A is None and bell.A_pattern.nnz == 322412


def test_unknown_formats_are_rejected(tmp_path):
    np.savez(tmp_path / "other.npz", format=np.asarray("something-else"), version=np.asarray(1))
    with pytest.raises(formats.FormatError, match="format"):
        formats.load_system(tmp_path / "other.npz")
    np.savez(tmp_path / "future.npz", format=np.asarray(formats.FORMAT), version=np.asarray(99),
             meta_json=np.asarray("{}"))
    with pytest.raises(formats.FormatError, match="version 99"):
        formats.load_system(tmp_path / "future.npz")
    np.savez(tmp_path / "bare.npz", x=np.zeros(3))
    with pytest.raises(formats.FormatError):
        formats.load_system(tmp_path / "bare.npz")
