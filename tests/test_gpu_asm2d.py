"""GPU 2D cut-cell assembly (``cc_cut2d_gram``, ``cc_asm2d_count``,
``cc_asm2d_fill`` through ``cutcell.asm2d_gpu``) against the host assembler
and against the fixtures the reference's own code wrote.

Bars: the CSR patterns of H and K identical to the host's
(``cutcell.oblique2d.assemble_oblique_cut``, which follows
``code/oblique_cut_sbp.py:271-433``); entries within 1e-12 of the largest
entry (the device evaluates the B-splines by Cox-de Boor and sums the
covering cells in another order); the counts of contributing cut cells and
the smallest cut fraction equal; and, on every committed 2D fixture, the
index maps, the E pattern and the A pattern of the reference-generated
system bit-exact when E and A are built from the device-assembled H and K.
"""

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from cutcell.basis import make_open_uniform_basis
from cutcell.oblique2d import HalfPlaneCut, assemble_oblique_cut, polynomial_preserving_extension, pullback, sym
from cutcell.workloads import c3_cases

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "agg_manifest.json").read_text())
CASES_2D = [k for k, v in MANIFEST["cases"].items() if v["kind"] == "oblique2d"]


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip

    hip.load_library()
    return hip


def _same(M, R):
    M, R = sp.csr_matrix(M), sp.csr_matrix(R)
    M.sort_indices()
    R.sort_indices()
    np.testing.assert_array_equal(M.indptr, R.indptr)
    np.testing.assert_array_equal(M.indices, R.indices)
    assert np.abs(M.data - R.data).max() <= 1e-12 * np.abs(R.data).max()


CUTS = {
    "C1-5deg": (32, 4, np.deg2rad(5.0), 0.85 / 32, 24),
    "sliver-eps1e-5": (16, 4, 0.0, (1.0 - 1e-5) / 16, 24),
    "steep-40deg-mid": (20, 3, np.deg2rad(40.0), 0.45, 20),
    "p5-60deg": (24, 5, np.deg2rad(60.0), 0.6, 24),
    "p1-20deg": (12, 1, np.deg2rad(20.0), 0.3, 10),
}


@pytest.mark.parametrize("name", sorted(CUTS))
def test_device_assembly_matches_host(gpu, name):
    from cutcell.asm2d_gpu import assemble_oblique_cut_gpu

    N, p, ang, off, q = CUTS[name]
    bx, by = make_open_uniform_basis(N, p), make_open_uniform_basis(N, p)
    cut = HalfPlaneCut.from_angle(ang, off)
    ref = assemble_oblique_cut(bx, by, cut, q)
    got = assemble_oblique_cut_gpu(bx, by, cut, q)
    print(f"{name}: {got.n_cut_cells} cut cells, nnz {got.H.nnz}", flush=True)
    _same(got.H, ref.H)
    _same(got.K, ref.K)
    assert got.n_cut_cells == ref.n_cut_cells
    assert got.min_cut_fraction == pytest.approx(ref.min_cut_fraction, rel=1e-12)
    np.testing.assert_array_equal(got.H_background_diag, ref.H_background_diag)


def test_batched_assembly_of_the_c3_cases(gpu):
    """The first 24 oblique cases of the batched configuration (p = 3, 4, 5
    on 48 x 48), each degree's cut cells in one launch, against the host."""
    from cutcell.asm2d_gpu import assemble_batch_gpu

    cases = [c for c in c3_cases() if c[0] == "oblique"][:24]
    for p in (3, 4, 5):
        sel = [c for c in cases if c[2] == p]
        basis = make_open_uniform_basis(48, p)
        cuts = [HalfPlaneCut.from_angle(np.deg2rad(c[3]), c[4] / 48) for c in sel]
        got = assemble_batch_gpu(basis, cuts, max(24, 2 * p + 14))
        for cut, g in zip(cuts, got):
            ref = assemble_oblique_cut(basis, basis, cut, max(24, 2 * p + 14))
            _same(g.H, ref.H)
            _same(g.K, ref.K)
            assert g.n_cut_cells == ref.n_cut_cells


@pytest.mark.parametrize("name", CASES_2D)
def test_reference_fixtures_from_device_assembly(gpu, name):
    from cutcell.asm2d_gpu import assemble_oblique_cut_gpu

    meta = MANIFEST["cases"][name]
    ref = dict(np.load(GOLDEN / f"agg_{name}.npz", allow_pickle=False))
    p, N = meta["degree"], meta["n_elements"]
    bx, by = make_open_uniform_basis(N, p), make_open_uniform_basis(N, p)
    s = assemble_oblique_cut_gpu(bx, by, HalfPlaneCut.from_angle(meta["angle"], meta["offset"]),
                                 meta["quadrature_order"])
    agg = polynomial_preserving_extension(s, meta["support_threshold"], p, active_tol=meta["active_tol"])
    np.testing.assert_array_equal(agg.active, ref["active"])
    np.testing.assert_array_equal(agg.interior, ref["interior"])
    np.testing.assert_array_equal(agg.small_support, ref["small"])
    E = sp.csr_matrix(agg.E)
    E.sort_indices()
    np.testing.assert_array_equal(E.indptr, ref["E_indptr"])
    np.testing.assert_array_equal(E.indices, ref["E_indices"])
    A = sp.csr_matrix(sym(pullback(s.K, agg.E) + pullback(s.H, agg.E)))
    A.sort_indices()
    np.testing.assert_array_equal(A.indptr, ref["A_indptr"])
    np.testing.assert_array_equal(A.indices, ref["A_indices"])
    assert np.abs(A.data - ref["A_data"]).max() <= 1e-11 * np.abs(ref["A_data"]).max()
