"""GPU parity of the HIP Jacobi-PCG path against the CPU oracle.

Every test calls the hand-written gfx950 kernels through the C ABI
(``cutcell.hip`` -> ``libcutcell_hip.so``) and compares with ``oracle`` on the
same inputs. Floating-point bars: SpMV to 1e-13 relative (float64); the
residual trajectory of the oracle's restatement of the reference GPU loop
(``parity.check_trajectory``: same envelope within 2x, stops at most
``parity.MAX_ITER_GAP`` = 3 iterations apart, and the earlier stop only where
the other run's residual is within ``parity.STOP_SLACK`` = 2x of tol);
solutions within 1e-6 of the oracle's at tol 1e-8; float64 true residual
<= 10 tol; and, at tol 1e-12, within 1e-8 of SuperLU.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from cutcell.oblique2d import oblique_system, sliver_system
from parity import check_stop_gap, check_trajectory

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip as h

    h.load_library()
    return h


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def systems():
    return {
        "sliver-8": sliver_system(8, 4, 1e-5),
        "sliver-16": sliver_system(16, 4, 1e-2),
        "sliver-32": sliver_system(32, 4, 1e-3),
        "C1-oblique-5deg": oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32),
    }


@pytest.mark.parametrize("name", ["sliver-8", "sliver-32", "C1-oblique-5deg"])
def test_spmv_f64_matches_oracle(hip, systems, name):
    A = systems[name].A
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    Ad = hip.DeviceCSR.from_scipy(A)
    y = hip.spmv(Ad, _dev(x)).cpu().numpy()
    ref = oracle.spmv(A, x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(A) @ np.abs(x))


def test_spmv_f32(hip, systems):
    A = systems["sliver-32"].A
    x = np.random.default_rng(2).standard_normal(A.shape[0]).astype(np.float32)
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    y = hip.spmv(Ad, _dev(x)).cpu().numpy()
    A32 = A.astype(np.float32).astype(np.float64)
    ref = A32 @ x.astype(np.float64)
    assert np.max(np.abs(y - ref)) <= 1e-5 * np.max(np.abs(A32) @ np.abs(x.astype(np.float64)))


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_spmm_columns(hip, systems, k):
    A = systems["sliver-16"].A
    X = np.random.default_rng(k).standard_normal((A.shape[0], k))
    Y = hip.spmm(hip.DeviceCSR.from_scipy(A), _dev(X)).cpu().numpy()
    ref = np.stack([oracle.spmv(A, X[:, c]) for c in range(k)], axis=1)
    np.testing.assert_allclose(Y, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def _ragged(n=700, seed=5):
    """Rows of length 0 to ~200 (empty rows included), random columns."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 200, n)
    lens[::37] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, rows.size)
    A = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    return A


def _long_rows(n=12000, seed=9):
    """Short rows, empty rows and a few rows longer than the window kernel's
    4096-entry window (their sums cross two or more windows)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n)
    lens[::53] = 0
    lens[[7, 2500, 2501, n - 1]] = [5000, 4096, 9000, 4500]
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([rng.choice(n, size=m, replace=False) for m in lens])
    A = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


@pytest.mark.parametrize("k,vb", [(1, 8), (1, 4), (16, 8), (16, 4), (4, 4)])
def test_spmm_every_launch_shape(hip, systems, k, vb):
    """Each launch shape the autotuner may pick (vector width, nonzero slots,
    grid, unroll 1/4/8, nontemporal loads, and for k = 1 the nonzero-window
    kernel) against the oracle SpMV, on a B-spline system, on a ragged matrix
    with empty rows and on one with rows longer than a window."""
    from cutcell.autotune import spmm_candidates

    lib = hip.load_library()
    mats = {"sliver-16": systems["sliver-16"].A, "ragged": _ragged(), "long-rows": _long_rows()}
    try:
        for name, A in mats.items():
            Ad = hip.DeviceCSR.from_scipy(A, with_f32=(vb == 4))
            X = np.random.default_rng(k).standard_normal((A.shape[0], k))
            if vb == 4:
                X = X.astype(np.float32)
            A_ref = A.astype(np.float32).astype(np.float64) if vb == 4 else A
            ref = np.stack([oracle.spmv(A_ref.tocsr(), X[:, c].astype(np.float64)) for c in range(k)], axis=1)
            scale = (np.abs(A_ref) @ np.abs(X.astype(np.float64))).max()
            tol = 1e-5 if vb == 4 else 1e-13
            seen = 0
            for cand in spmm_candidates(k, A.nnz / A.shape[0], vb):
                if lib.cc_spmm_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"],
                                          cand["stream"], cand["sweep"]) != 0:
                    continue
                Y = hip.spmm(Ad, _dev(X)).cpu().numpy() if k > 1 else hip.spmv(Ad, _dev(X[:, 0])).cpu().numpy()[:, None]
                assert np.abs(Y - ref).max() <= tol * scale, (name, cand)
                seen += 1
            assert seen >= 5
    finally:
        lib.cc_spmm_set_config(k, vb, 0, 0, 0, 0, 0, 0)


@pytest.mark.parametrize("stream,sweep", [(0, 1), (1, 8)])
def test_window_spmv_in_pcg_and_residual(hip, systems, stream, sweep):
    """The nonzero-window kernel as the product of the single-right-hand-side
    CG loop (with the fused p.Ap partials) and of the float64 true residual:
    the same trajectory bar as the row kernels, and the device residual equal
    to the host one."""
    from cutcell.autotune import WINDOW_UNROLL

    lib = hip.load_library()
    s = systems["C1-oblique-5deg"]
    Ad = hip.DeviceCSR.from_scipy(s.A)
    try:
        assert lib.cc_spmm_set_config(1, 8, 64, 0, WINDOW_UNROLL, 1, stream, sweep) == 0
        xd, res, hist = hip.pcg(Ad, _dev(s.b), tol=1e-8, maxiter=5000, history=True)
        r = res[0]
        x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=1e-8, history=True)
        assert r.status == hip.CC_CONVERGED == st_or
        check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
        x = xd.cpu().numpy()
        assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < 1e-6
        assert r.true_rel == pytest.approx(np.linalg.norm(s.A @ x - s.b) / np.linalg.norm(s.b), rel=1e-6)
        X = np.random.default_rng(3).standard_normal(s.A.shape[0])
        rel = hip.residual(Ad, _dev(s.b), _dev(X))[0]
        assert rel == pytest.approx(np.linalg.norm(s.b - s.A @ X) / np.linalg.norm(s.b), rel=1e-12)
    finally:
        lib.cc_spmm_set_config(1, 8, 0, 0, 0, 0, 0, 0)


def test_jacobi_diag(hip, systems):
    A = systems["C1-oblique-5deg"].A
    d = hip.jacobi_diag(hip.DeviceCSR.from_scipy(A)).cpu().numpy()
    np.testing.assert_array_equal(d, A.diagonal())


@pytest.mark.parametrize("name", ["sliver-8", "sliver-16", "sliver-32", "C1-oblique-5deg"])
def test_pcg_matches_oracle(hip, systems, name):
    s = systems[name]
    Ad = hip.DeviceCSR.from_scipy(s.A)
    xd, res, hist = hip.pcg(Ad, _dev(s.b), tol=1e-8, maxiter=5000, history=True)
    r = res[0]
    x_or, it_or, st_or, rel_or, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=1e-8, history=True)
    print(f"{name}: gpu it={r.iterations} status={r.status} true_rel={r.true_rel:.3e}; oracle it={it_or}")
    assert r.status == hip.CC_CONVERGED == st_or
    check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
    assert hist[r.iterations, 0] == pytest.approx(r.recursive_rel, rel=1e-12)
    assert r.true_rel <= 1e-7
    x = xd.cpu().numpy()
    host_rel = np.linalg.norm(s.A @ x - s.b) / np.linalg.norm(s.b)
    assert r.true_rel == pytest.approx(host_rel, rel=1e-6)
    assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < 1e-6


@pytest.mark.parametrize("name", ["sliver-16", "C1-oblique-5deg"])
def test_pcg_solution_against_superlu(hip, systems, name):
    s = systems[name]
    xd, res = hip.pcg(hip.DeviceCSR.from_scipy(s.A), _dev(s.b), tol=1e-12, maxiter=5000)
    xs = spla.splu(s.A.tocsc()).solve(s.b)
    x = xd.cpu().numpy()
    assert res[0].status == hip.CC_CONVERGED
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


@pytest.mark.parametrize("k", [2, 16])
def test_multi_rhs_matches_single_rhs(hip, systems, k):
    s = systems["sliver-16"]
    rng = np.random.default_rng(7)
    B = rng.standard_normal((s.A.shape[0], k))
    B[:, 0] = s.b
    Ad = hip.DeviceCSR.from_scipy(s.A)
    X, res, hist = hip.pcg(Ad, _dev(B), tol=1e-8, history=True)
    X1, res1, hist1 = hip.pcg(Ad, _dev(np.ascontiguousarray(B[:, 0])), tol=1e-8, history=True)
    X = X.cpu().numpy()
    check_trajectory(hist[:, 0], res[0].iterations, hist1[:, 0], res1[0].iterations, 1e-8)
    for c in range(k):
        x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, B[:, c], tol=1e-8, history=True)
        assert res[c].status == hip.CC_CONVERGED
        check_trajectory(hist[:, c], res[c].iterations, h_or, it_or, 1e-8)
        assert np.linalg.norm(X[:, c] - x_or) / np.linalg.norm(x_or) < 1e-6
        assert res[c].true_rel <= 1e-7


@pytest.mark.parametrize("k", [1, 16])
def test_mixed_precision_reaches_fp64_tolerance(hip, systems, k):
    s = systems["sliver-32"]
    B = np.random.default_rng(11).standard_normal((s.A.shape[0], k))
    Ad = hip.DeviceCSR.from_scipy(s.A, with_f32=True)
    X, res = hip.pcg(Ad, _dev(B), tol=1e-8, precision="mixed", inner_tol=1e-4)
    X = X.cpu().numpy()
    lu = spla.splu(s.A.tocsc())
    for c in range(k):
        rel = np.linalg.norm(s.A @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
        assert res[c].status == hip.CC_CONVERGED and rel <= 1e-8
        assert res[c].true_rel == pytest.approx(rel, rel=1e-6)
        xs = lu.solve(B[:, c])
        assert np.linalg.norm(X[:, c] - xs) / np.linalg.norm(xs) < 1e-5
    _, total, sweeps, st, _ = oracle.mixed_refinement_py(s.A, B[:, 0], tol=1e-8, inner_tol=1e-4)
    assert abs(res[0].outer - sweeps) <= 1
    assert abs(res[0].inner_total - total) <= 0.1 * total + 3


def test_edge_cases(hip):
    n = 7
    I = sp.eye(n, format="csr")
    # zero right-hand side: the reference GPU loop breaks down at k = 1
    _, res = hip.pcg(hip.DeviceCSR.from_scipy(I), _dev(np.zeros(n)))
    assert (res[0].iterations, res[0].status) == (1, hip.CC_BREAKDOWN)
    # negative definite
    _, res = hip.pcg(hip.DeviceCSR.from_scipy(-I), _dev(np.ones(n)))
    assert (res[0].iterations, res[0].status) == (1, hip.CC_BREAKDOWN)
    # maxiter exhaustion
    s = sliver_system(8, 4, 1e-3)
    _, res = hip.pcg(hip.DeviceCSR.from_scipy(s.A), _dev(s.b), maxiter=3)
    assert (res[0].iterations, res[0].status) == (3, hip.CC_MAXITER)
    # zero diagonal -> identity (hand-computed: breakdown at k = 2, x = [3, 1.5])
    A = sp.csr_matrix(np.array([[0.0, 0.0], [0.0, 2.0]]))
    x, res = hip.pcg(hip.DeviceCSR.from_scipy(A), _dev(np.ones(2)), maxiter=5)
    assert (res[0].iterations, res[0].status) == (2, hip.CC_BREAKDOWN)
    np.testing.assert_array_equal(x.cpu().numpy(), [3.0, 1.5])
    # one unknown
    x, res = hip.pcg(hip.DeviceCSR.from_scipy(sp.csr_matrix([[4.0]])), _dev(np.array([2.0])))
    assert res[0].status == hip.CC_CONVERGED and x.cpu().item() == 0.5


def test_solve_spd_gpu_report(hip, systems):
    from fast_solver_backends import solve_spd

    A = sp.diags([2.0, 3.0, 4.0, 5.0], format="csr")
    b = np.array([1.0, -2.0, 0.5, 4.0])
    x, rep = solve_spd(A, b, backend="gpu", tol=1e-12, maxiter=20)
    assert rep.backend == "gpu" and rep.method == "hip-jacobi-pcg" and rep.success, rep
    assert rep.iterations == 1 and rep.relative_residual < 1e-10
    np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-12)
    s = systems["C1-oblique-5deg"]
    x, rep = solve_spd(s.A, s.b, backend="gpu")
    _, it_or, _, _ = oracle.pcg_gpu_loop(s.A, s.b)
    assert rep.success and rep.transfer_s > 0
    check_stop_gap(rep.iterations, it_or)
    x, rep = solve_spd(s.A, s.b, backend="auto")
    assert rep.backend == "gpu" and rep.success
