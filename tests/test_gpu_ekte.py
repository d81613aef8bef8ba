"""GPU parity of the E-fused operator (``cc_ekte_apply``, ``cc_pcg_ekte``)
against the reference's formed ``sym(E^T K E) + sym(E^T H E)``
(``code/bell_3d_core.py:550-552``, built on the host by
``cutcell.cut3d.restrict_spd``).

Bars: the apply within 1e-11 of ``|A| |x|`` (the formed and fused operators
sum the same products in different orders); k columns equal k single
applies to the same bar; the solve within ``parity.MAX_ITER_GAP`` of the oracle's stop on the
formed matrix (``parity.check_stop_gap``), solutions within 1e-6 of SuperLU's; on
the reference's published bell system N = 20 the iterations within
``parity.REFERENCE_ITER_GAP`` of the published 104.
"""

import csv
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from parity import REFERENCE_ITER_GAP, check_stop_gap

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module", params=[False, True], ids=["csr", "row-blocks"])
def planar(request):
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip
    from cutcell.cut3d import PlanarCut, polynomial_preserving_extension_3d, restrict_spd, assemble_levelset_3d
    from cutcell.ekte import DeviceEkte, ekte_parts

    hip.load_library()
    cut = PlanarCut.random(0)
    asm = assemble_levelset_3d(cut, 10, 3)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    parts = ekte_parts(asm.K + asm.H, agg.E, agg.interior, agg.small_support)
    print(f"planar N=10 p=3: n={A.shape[0]} small={parts.n_small} touched={parts.t_rows.size} "
          f"nnz(A)={A.nnz} nnz(B)={parts.B.nnz}", flush=True)
    return A, parts, DeviceEkte(parts, row_blocks=request.param)


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_apply_matches_formed_operator(planar, k):
    A, parts, dev = planar
    X = np.random.default_rng(k).standard_normal((A.shape[0], k))
    Y = dev.apply(_dev(X)).cpu().numpy()
    scale = (abs(A) @ np.abs(X)).max()
    assert np.abs(Y - A @ X).max() <= 1e-11 * scale
    if k > 1:
        Y1 = np.stack([dev.apply(_dev(X[:, c])).cpu().numpy() for c in range(k)], axis=1)
        assert np.abs(Y - Y1).max() <= 1e-12 * scale


def test_pcg_on_fused_operator_matches_oracle_and_superlu(planar):
    A, parts, dev = planar
    from cutcell import hip

    b = np.random.default_rng(5).standard_normal(A.shape[0])
    x, res = dev.pcg(_dev(b), tol=1e-8, maxiter=5000)
    r = res[0]
    x_or, it_or, st_or, _ = oracle.pcg_gpu_loop(A, b, tol=1e-8)
    assert r.status == hip.CC_CONVERGED == st_or
    check_stop_gap(r.iterations, it_or)
    x = x.cpu().numpy()
    rel = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    assert r.true_rel == pytest.approx(rel, rel=1e-5) and rel <= 1e-7
    xs = spla.splu(A.tocsc()).solve(b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    # 16 right-hand sides in one solve: each column as its single solve.
    B = np.random.default_rng(6).standard_normal((A.shape[0], 16))
    B[:, 0] = b
    X, resk = dev.pcg(_dev(B), tol=1e-8, maxiter=5000)
    X = X.cpu().numpy()
    assert all(rr.status == hip.CC_CONVERGED for rr in resk)
    check_stop_gap(resk[0].iterations, r.iterations)
    for c in range(16):
        assert np.linalg.norm(A @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-7


def test_fused_operator_without_small_dofs():
    """No small dof (E the identity): the fused apply is the plain product
    with B and launches no gather or scatter."""
    from cutcell.ekte import DeviceEkte, ekte_parts

    n = 50
    rng = np.random.default_rng(2)
    R = sp.random(n, n, density=0.1, random_state=3, format="csr")
    Bm = (R @ R.T + sp.identity(n)).tocsr()
    parts = ekte_parts(Bm, sp.identity(n, format="csr"), np.arange(n), np.zeros(0, dtype=np.int64))
    assert parts.n_small == 0 and parts.t_rows.size == 0
    dev = DeviceEkte(parts)
    x = rng.standard_normal(n)
    y = dev.apply(_dev(x)).cpu().numpy()
    Bs = 0.5 * (Bm + Bm.T)
    assert np.abs(y - Bs @ x).max() <= 1e-13 * (abs(Bs) @ np.abs(x)).max()


def test_bell_published_system_on_fused_operator():
    """The reference's published bell system N = 20 (p = 2, eps = 1e-3), one
    right-hand side b = E^T load, solved on the fused operator: iterations
    within REFERENCE_ITER_GAP of the reference GPU loop's 104, the float64
    true residual of the formed system below 1e-8."""
    from cutcell import hip
    from cutcell.bell3d import assemble_bell
    from cutcell.cut3d import polynomial_preserving_extension_3d, restrict_spd
    from cutcell.ekte import DeviceEkte, ekte_parts

    row = next(csv.DictReader((GOLDEN / "reference_gpu_solver_benchmark_3d.csv").open()))
    assert int(row["n_elements"]) == 20
    asm, load = assemble_bell(20, 2)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    b = np.asarray(agg.E.T @ load)
    dev = DeviceEkte(ekte_parts(asm.K + asm.H, agg.E, agg.interior, agg.small_support))
    x, res = dev.pcg(_dev(b), tol=1e-8, maxiter=20000)
    r = res[0]
    _, it_or, st_or, _ = oracle.pcg_gpu_loop(A, b, tol=1e-8, maxiter=20000)
    print(f"bell N=20 fused: HIP {r.iterations}, oracle (formed) {it_or}, reference {row['gpu_cg_iterations']}",
          flush=True)
    assert r.status == hip.CC_CONVERGED == st_or
    assert abs(r.iterations - int(row["gpu_cg_iterations"])) <= REFERENCE_ITER_GAP
    x = x.cpu().numpy()
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-7
