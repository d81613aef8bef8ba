"""GPU checks at the benchmark workloads' full sizes, through properties that
do not depend on size (the oracle cannot run these sizes in seconds):

- the 64^3 sphere box (n = 266,234) and the 128x64x64 bump channel
  (n = 508,672), matrix-free with their cut cells integrated on the device:
  the operator is symmetric and positive on random vectors, the 16-column
  apply equals 16 one-column applies, and the channel's pressure operator
  E^T K E maps constants to zero (its null space);
- the network pass of the learned preconditioner on the channel's five grids
  and 16 columns, where the blocks of all levels must fit the resident slots:
  the three-launch pass against the per-level layer launches;
- the 3D planar-cut system (N = 43, n = 81,195, 27.4M nonzeros): the row-block
  product against the CSR product, and the headline solve (float32 row-block
  products in float64 refinement, 16 right-hand sides) reaching the float64
  true residual 1e-8 (recomputed on the host) with solutions within 1e-6 of
  the float64 solve's.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip as h

    h.require_gpu()
    return h


@pytest.fixture(scope="module")
def sphere(hip):
    from cutcell.matfree import DeviceMatrixFree
    from cutcell.workloads import c4_operator

    mf = c4_operator(0, n_elements=64, degree=3, device=True)
    print(f"sphere box: n = {mf.n}", flush=True)
    return mf, DeviceMatrixFree(mf)


@pytest.fixture(scope="module")
def channel(hip):
    from cutcell.matfree import DeviceMatrixFree
    from cutcell.workloads import c5_operator

    mf = c5_operator(mass_coef=0.0, device=True)
    print(f"bump channel: n = {mf.n}", flush=True)
    return mf, DeviceMatrixFree(mf)


def _operator_properties(dmf, n, rng):
    """Symmetry and positivity on random columns, and the 16-column apply
    against one-column applies (different kernels, same sums up to order)."""
    X = _dev(rng.standard_normal((n, 16)))
    Y = dmf.apply_k(X)
    torch.cuda.synchronize()
    Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
    assert np.isfinite(Yh).all()
    for a, b in [(0, 1), (2, 9), (5, 15)]:
        lhs, rhs = Xh[:, a] @ Yh[:, b], Xh[:, b] @ Yh[:, a]
        scale = np.linalg.norm(Xh[:, a]) * np.linalg.norm(Yh[:, b])
        assert abs(lhs - rhs) <= 1e-12 * scale, (a, b, lhs, rhs)
    assert (np.einsum("ij,ij->j", Xh, Yh) > 0).all()
    for c in (0, 7, 15):
        y1 = dmf.apply(X[:, c].contiguous()).cpu().numpy()
        err = np.linalg.norm(y1 - Yh[:, c]) / np.linalg.norm(Yh[:, c])
        assert err <= 1e-12, (c, err)
    return Xh, Yh


def test_sphere_operator_at_full_size(sphere):
    mf, dmf = sphere
    assert mf.n == 266_234
    _operator_properties(dmf, mf.n, np.random.default_rng(3))


def test_channel_pressure_operator_at_full_size(channel):
    """Besides the common properties: A 1 = 0 to rounding, measured against
    the operator's scale ||A x|| / ||x|| on random x."""
    mf, dmf = channel
    assert mf.n == 508_672
    Xh, Yh = _operator_properties(dmf, mf.n, np.random.default_rng(4))
    scale = np.linalg.norm(Yh[:, 0]) / np.linalg.norm(Xh[:, 0])
    ones = torch.ones(mf.n, dtype=torch.float64, device="cuda")
    r = dmf.apply(ones).cpu().numpy()
    ratio = np.linalg.norm(r) / np.sqrt(mf.n) / scale
    print(f"channel: ||A 1|| / ||1|| = {ratio:.2e} of ||A x|| / ||x||", flush=True)
    assert ratio <= 1e-10


def test_network_pass_at_full_channel_size(channel):
    """One flexible-CG step (x_1 = alpha z_0, z_0 the preconditioner of r_0)
    on 16 columns with an untrained five-level network: the three-launch pass
    (run lengths grown until all levels' blocks are resident) against the
    per-level layer launches, both bf16 with the same rounding points."""
    from cutcell import learned
    from cutcell.learned import MultilevelCNN

    mf, dmf = channel
    periodic = tuple(a.periodic for a in mf.grid.axes)
    net = MultilevelCNN(mf.grid.nb, mf.interior, periodic=periodic, levels=5, seed=0)
    lib = learned._lib()
    B = _dev(np.random.default_rng(5).standard_normal((mf.n, 16)))
    B = B - B.mean(dim=0, keepdim=True)  # compatible with the constants' null space
    B = B.contiguous()
    try:
        assert lib.cc_net_set_config(0, 1) == 0
        X_old, _ = net.fcg_batched(dmf, B, tol=0.0, maxiter=1)
        X_old = X_old.cpu().numpy()
        assert lib.cc_net_set_config(0, 0) == 0
        X_new, _ = net.fcg_batched(dmf, B, tol=0.0, maxiter=1)
        X_new = X_new.cpu().numpy()
    finally:
        lib.cc_net_set_config(0, 0)
    assert np.isfinite(X_new).all()
    err = np.linalg.norm(X_new - X_old, axis=0) / np.linalg.norm(X_old, axis=0)
    print(f"channel network pass: three-launch vs per-level, worst column rel diff {err.max():.2e}", flush=True)
    assert err.max() <= 2e-3


@pytest.fixture(scope="module")
def planar(hip):
    from cutcell.workloads import c2_system

    s = c2_system(seed=0, target=81_000)
    print(f"planar system: n = {s.A.shape[0]}, nnz = {s.A.nnz}", flush=True)
    return s


def test_planar_products_at_full_size(hip, planar):
    """Row-block and CSR products of the 27.4M-nonzero matrix on 16 columns
    agree to rounding (float64), and x.Ay = y.Ax."""
    A = planar.A
    assert (A.shape[0], A.nnz) == (81_195, 27_369_621)
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    X = _dev(np.random.default_rng(6).standard_normal((A.shape[0], 16)))
    Y_csr = hip.spmm(Ad, X).cpu().numpy()
    Y_rb = hip.spmm_rb(rb, X).cpu().numpy()
    Xh = X.cpu().numpy()
    err = np.linalg.norm(Y_rb - Y_csr, axis=0) / np.linalg.norm(Y_csr, axis=0)
    assert err.max() <= 1e-13, err.max()
    lhs, rhs = Xh[:, 0] @ Y_rb[:, 1], Xh[:, 1] @ Y_rb[:, 0]
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Xh[:, 0]) * np.linalg.norm(Y_rb[:, 1])


def test_planar_headline_solve_at_full_size(hip, planar):
    A = planar.A
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    B = np.random.default_rng(7).standard_normal((A.shape[0], 16))
    Xm, res_m = hip.pcg(Ad, _dev(B), tol=1e-8, precision="mixed", inner_tol=1e-4, rb=rb)
    X64, res_64 = hip.pcg(Ad, _dev(B), tol=1e-8, precision="fp64", rb=rb)
    Xm, X64 = Xm.cpu().numpy(), X64.cpu().numpy()
    for c in range(16):
        rel = np.linalg.norm(A @ Xm[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
        assert res_m[c].status == hip.CC_CONVERGED and rel <= 1e-8, (c, rel)
        assert res_m[c].true_rel == pytest.approx(rel, rel=1e-5)
        assert res_64[c].status == hip.CC_CONVERGED
        diff = np.linalg.norm(Xm[:, c] - X64[:, c]) / np.linalg.norm(X64[:, c])
        assert diff <= 1e-6, (c, diff)
    print(f"planar headline: inner iterations {min(r.inner_total for r in res_m)}-"
          f"{max(r.inner_total for r in res_m)}, sweeps {max(r.outer for r in res_m)}", flush=True)
