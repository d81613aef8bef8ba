"""Aggregation parity against fixtures written by the reference's own code.

``tests/golden/make_golden.py`` ran the reference's
``assemble_oblique_cut_matrices`` + ``polynomial_preserving_oblique_extension``
(``code/oblique_cut_sbp.py:271-433, 614-672``) and ``assemble_bell_3d`` +
``polynomial_preserving_extension_3d`` + ``restrict_spd``
(``code/bell_3d_core.py:275-430, 471-552``) and stored their outputs. Here the
system builders are run on the same inputs and compared:

- integer outputs bit-exact: the active, interior and small index maps, the
  CSR pattern of ``E`` and of the aggregated operator ``A`` (its nnz set by
  exact-zero cancellations);
- float outputs to rounding: ``gamma``, the entries of ``E`` and ``A``.

The 3D rule sorts Greville distances with NumPy's default (unstable)
``argsort``, whose tie order depends on the host's NumPy build and SIMD
dispatch; the fixture records the container it came from, and the builder's
``tie_order="numpy-default"`` mode reproduces it there. The builder's default
``"stable"`` order (ties by dof index) must give the same index maps and
row supports of the same size.
"""

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from cutcell.basis import make_open_uniform_basis
from cutcell.bell3d import assemble_bell
from cutcell.cut3d import polynomial_preserving_extension_3d, restrict_spd
from cutcell.oblique2d import HalfPlaneCut, assemble_oblique_cut, polynomial_preserving_extension, pullback, sym

GOLDEN = Path(__file__).resolve().parent / "golden"
MANIFEST = json.loads((GOLDEN / "agg_manifest.json").read_text())
CASES = MANIFEST["cases"]
CASES_2D = [k for k, v in CASES.items() if v["kind"] == "oblique2d"]
CASES_3D = [k for k, v in CASES.items() if v["kind"] == "bell3d"]


def _load(name):
    return dict(np.load(GOLDEN / f"agg_{name}.npz", allow_pickle=False))


def _csr(d, prefix):
    shape = tuple(int(v) for v in d[f"{prefix}_shape"])
    data = d.get(f"{prefix}_data")
    if data is None:
        data = np.ones(d[f"{prefix}_indices"].size)
    return sp.csr_matrix((data, d[f"{prefix}_indices"], d[f"{prefix}_indptr"]), shape=shape)


def _same_pattern(M, d, prefix):
    M = sp.csr_matrix(M)
    M.sort_indices()
    assert M.shape == tuple(int(v) for v in d[f"{prefix}_shape"])
    np.testing.assert_array_equal(M.indptr, d[f"{prefix}_indptr"])
    np.testing.assert_array_equal(M.indices, d[f"{prefix}_indices"])
    return M


@pytest.mark.parametrize("name", CASES_2D)
def test_2d_index_maps_and_e_pattern_bit_exact(name):
    meta, ref = CASES[name], _load(name)
    p, N = meta["degree"], meta["n_elements"]
    bx = make_open_uniform_basis(N, p)
    by = make_open_uniform_basis(N, p)
    s = assemble_oblique_cut(bx, by, HalfPlaneCut.from_angle(meta["angle"], meta["offset"]),
                             quadrature_order=meta["quadrature_order"])
    agg = polynomial_preserving_extension(s, meta["support_threshold"], p, active_tol=meta["active_tol"])
    np.testing.assert_array_equal(agg.active, ref["active"])
    np.testing.assert_array_equal(agg.interior, ref["interior"])
    np.testing.assert_array_equal(agg.small_support, ref["small"])
    np.testing.assert_allclose(agg.gamma, ref["gamma"], rtol=1e-12, atol=1e-300)
    assert agg.polynomial_degree == int(ref["polynomial_degree"])
    E = _same_pattern(agg.E, ref, "E")
    np.testing.assert_allclose(E.data, ref["E_data"], rtol=1e-9, atol=1e-12)
    A = sym(pullback(s.K, agg.E) + pullback(s.H, agg.E))
    A = _same_pattern(A, ref, "A")
    scale = np.abs(ref["A_data"]).max()
    assert np.abs(A.data - ref["A_data"]).max() <= 1e-11 * scale


def _bell(meta, tie_order):
    asm, _ = assemble_bell(meta["n_elements"], meta["degree"], meta["epsilon"])
    agg = polynomial_preserving_extension_3d(asm, tie_order=tie_order)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    return asm, agg, A


@pytest.mark.parametrize("name", CASES_3D)
def test_3d_index_maps_and_patterns_bit_exact(name):
    meta, ref = CASES[name], _load(name)
    asm, agg, A = _bell(meta, "numpy-default")
    np.testing.assert_array_equal(agg.active, ref["active"])
    np.testing.assert_array_equal(agg.interior, ref["interior"])
    np.testing.assert_array_equal(agg.small_support, ref["small"])
    np.testing.assert_allclose(agg.gamma, ref["gamma"], rtol=1e-12, atol=1e-300)
    E = _same_pattern(agg.E, ref, "E")
    np.testing.assert_allclose(E.data, ref["E_data"], rtol=1e-8, atol=1e-12)
    A.sort_indices()
    assert A.nnz == int(ref["A_nnz"])
    np.testing.assert_array_equal(np.diff(A.indptr), ref["A_row_nnz"])
    if "A_indices" in ref:
        _same_pattern(A, ref, "A")


@pytest.mark.parametrize("name", CASES_3D[:1])
def test_3d_stable_tie_order_keeps_maps_and_support_sizes(name):
    meta, ref = CASES[name], _load(name)
    _, agg, _ = _bell(meta, "stable")
    np.testing.assert_array_equal(agg.interior, ref["interior"])
    np.testing.assert_array_equal(agg.small_support, ref["small"])
    E_ref = _csr(ref, "E")
    np.testing.assert_array_equal(np.diff(sp.csr_matrix(agg.E).indptr), np.diff(E_ref.indptr))
    assert agg.polynomial_residual < 1e-10
