"""Host-side parts of the E-fused operator (``cutcell.ekte``): the factors
reproduce the reference's formed operator ``sym(E^T K E) + sym(E^T H E)``
(``code/bell_3d_core.py:550-552``) and its diagonal, on a 3D planar cut and on
the bell geometry, and the three-stage product the HIP kernels run (gather
u_s = E_s x, product with the renumbered B, scatter through T), emulated
with SciPy, gives the same y."""

import numpy as np
import pytest

from cutcell.cut3d import PlanarCut, build_system_3d
from cutcell.ekte import ekte_parts, ekte_parts_3d


def _three_stage(parts, x):
    n = parts.n_red
    u = parts.Es @ x
    yv = parts.B @ np.concatenate([x, u])
    y = yv[:n].copy()
    y[parts.t_rows] += parts.T @ yv[n:]
    return y


@pytest.fixture(scope="module")
def planar():
    cut = PlanarCut.random(3)
    sysm, asm, agg = build_system_3d(cut, 7, 3)
    parts, _, _ = ekte_parts_3d(cut, 7, 3)
    return sysm, parts


def test_parts_reproduce_the_formed_operator(planar):
    sysm, parts = planar
    A = sysm.A
    assert parts.n_red == A.shape[0] and parts.n_small > 0 and parts.t_rows.size > 0
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    y = _three_stage(parts, x)
    scale = (abs(A) @ np.abs(x)).max()
    assert np.abs(y - A @ x).max() <= 1e-12 * scale
    np.testing.assert_allclose(parts.diag, A.diagonal(), rtol=1e-12, atol=1e-14 * np.abs(A.diagonal()).max())
    F = parts.to_scipy_operator()
    assert abs(F - A).max() <= 1e-12 * abs(A).max()
    # Interior-only couplings are B's own entries; the fused form reads fewer
    # matrix entries than the formed one.
    assert parts.B.nnz < A.nnz * 1.2


def test_bell_parts_match_restrict_spd():
    from cutcell.bell3d import assemble_bell
    from cutcell.cut3d import polynomial_preserving_extension_3d, restrict_spd

    asm, _ = assemble_bell(20, 2)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    parts = ekte_parts(asm.K + asm.H, agg.E, agg.interior, agg.small_support)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    scale = (abs(A) @ np.abs(x)).max()
    assert np.abs(_three_stage(parts, x) - A @ x).max() <= 1e-12 * scale
    np.testing.assert_allclose(parts.diag, A.diagonal(), rtol=1e-11)
    assert parts.algorithmic_bytes(1) < A.nnz * 12 + 4 * (A.shape[0] + 1) + 16 * A.shape[0]
