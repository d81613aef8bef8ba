"""GPU checks of the learned preconditioner's kernels against the float64
NumPy oracle (oracle/cnn_ref.py) and of the flexible-CG solve."""

import ctypes

import numpy as np
import pytest

import oracle
from oracle import cnn_ref

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import learned

    return learned._lib()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_mfma_lane_layout(lib):
    from cutcell.learned import from_bf16_bits, to_bf16_bits

    rng = np.random.default_rng(0)
    A = to_bf16_bits(rng.standard_normal((16, 16)))
    B = to_bf16_bits(rng.standard_normal((16, 16)))
    D = torch.zeros(256, dtype=torch.float32, device="cuda")
    from cutcell import hip

    Ad, Bd = _dev(A), _dev(B)
    assert lib.cc_mfma_probe(Ad.data_ptr(), Bd.data_ptr(), D.data_ptr(), hip._stream()) == 0
    ref = from_bf16_bits(A).astype(np.float64) @ from_bf16_bits(B).astype(np.float64)
    np.testing.assert_allclose(D.cpu().numpy().reshape(16, 16), ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("periodic", [(False, False, False), (True, False, True)])
@pytest.mark.parametrize("relu", [False, True])
def test_conv_forward_matches_oracle(lib, periodic, relu):
    from cutcell import hip
    from cutcell.learned import from_bf16_bits, to_bf16_bits

    rng = np.random.default_rng(3)
    nz, ny, nx = 5, 6, 19
    x = from_bf16_bits(to_bf16_bits(rng.standard_normal((nz, ny, nx, 16)))).astype(np.float64)
    W = from_bf16_bits(to_bf16_bits(0.2 * rng.standard_normal((27, 16, 16)))).astype(np.float64)
    mask = (rng.random((nz, ny, nx)) > 0.25).astype(np.uint8)
    ref, _ = cnn_ref.conv3(x, W, mask.astype(float), relu=relu, periodic=periodic)
    out = torch.zeros(nz * ny * nx * 16, dtype=torch.float32, device="cuda")
    per = (ctypes.c_int32 * 3)(*[int(v) for v in periodic])
    # Device copies stay referenced until the kernel has run.
    xd, wd, md = _dev(to_bf16_bits(x)), _dev(to_bf16_bits(W)), _dev(mask)
    st = lib.cc_cnn_conv3_forward(nx, ny, nz, per, xd.data_ptr(), wd.data_ptr(), md.data_ptr(), None, int(relu), None,
                                  out.data_ptr(), hip._stream())
    torch.cuda.synchronize()
    assert st == 0
    got = out.cpu().numpy().reshape(nz, ny, nx, 16)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_network_gradient_matches_oracle(lib):
    """Loss and weight gradients of the whole network (bf16 activations on
    the GPU, float64 oracle): same direction, close magnitude."""
    from cutcell.learned import MaskedCNN, from_bf16_bits

    rng = np.random.default_rng(5)
    nz, ny, nx = 6, 7, 18
    mask = (rng.random((nz, ny, nx)) > 0.2)
    red2full = np.nonzero(mask.ravel())[0].astype(np.int64)
    net = MaskedCNN((nx, ny, nz), red2full, seed=1, init_scale=1.0)
    # Make the last layer non-trivial so every layer's gradient is exercised.
    w32 = net.w32.cpu().numpy()
    w32[3] *= 20.0
    net.w32.copy_(torch.from_numpy(w32))
    from cutcell.learned import to_bf16_bits

    net.wbf.copy_(torch.from_numpy(to_bf16_bits(w32).reshape(4, -1).astype(np.int16).view(np.uint16)))
    Ws = [from_bf16_bits(net.wbf[l].cpu().numpy()).astype(np.float64).reshape(27, 16, 16) for l in range(4)]
    r = rng.standard_normal(red2full.size)
    e = rng.standard_normal(red2full.size)
    rg = np.zeros(mask.size)
    rg[red2full] = r
    eg = np.zeros(mask.size)
    eg[red2full] = e
    L_ref, g_ref = cnn_ref.loss_and_grads(rg.reshape(nz, ny, nx), eg.reshape(nz, ny, nx), Ws, mask.astype(float))
    diag = torch.ones(red2full.size, dtype=torch.float64, device="cuda")
    L = net.sample_loss_grad(_dev(r.astype(np.float32)), _dev(e.astype(np.float32)), diag, 1.0, 1.0, False,
                             float(np.sum(e.astype(np.float32).astype(np.float64) ** 2)))
    assert L == pytest.approx(L_ref, rel=2e-2)
    got = net.grad.cpu().numpy().reshape(4, 27, 16, 16)
    for l in range(4):
        a, b = got[l].ravel(), g_ref[l].ravel()
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert cos > 0.99, (l, cos)
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b), l


def test_learned_fcg_trains_and_solves(lib):
    """Pairs from the solver's own iterates, a few Adam steps (loss falls),
    flexible CG to 1e-8 with the float64 true residual as the judge."""
    import scipy.sparse.linalg as spla

    from cutcell.cut3d import SphereObstacle
    from cutcell.grid3d import Axis, Grid3D
    from cutcell.learned import MaskedCNN, generate_pairs, solve_learned_with_fallback
    from cutcell.matfree import DeviceMatrixFree, assemble_reduced_csr, build_matrix_free

    grid = Grid3D(tuple(Axis(12, 3) for _ in range(3)))
    mf = build_matrix_free(grid, SphereObstacle.random(4))
    dmf = DeviceMatrixFree(mf)
    rng = np.random.default_rng(0)
    rhs = [_dev(rng.standard_normal(mf.n)) for _ in range(3)]
    R, E = generate_pairs(dmf, rhs, snapshots=[5, 15, 30, 60])
    assert R.shape == (12, mf.n) and torch.isfinite(R).all() and torch.isfinite(E).all()
    net = MaskedCNN(grid.nb, mf.interior, seed=0)
    diag = _dev(mf.diag)
    rep = net.train(R, E, diag, steps=40, batch=4, lr=3e-3)
    assert rep.loss_last < rep.loss_first
    b = rhs[0]
    x, res, hist = net.fcg(dmf, b, tol=1e-8, maxiter=2000, history=True)
    A = assemble_reduced_csr(mf)
    xs = spla.splu(A.tocsc()).solve(b.cpu().numpy())
    assert res.status == 1 and res.true_rel <= 1e-7
    xh = x.cpu().numpy()
    assert np.linalg.norm(A @ xh - b.cpu().numpy()) / np.linalg.norm(b.cpu().numpy()) == pytest.approx(res.true_rel, rel=1e-5)
    assert np.linalg.norm(xh - xs) / np.linalg.norm(xs) < 1e-5
    x2, res2, how = solve_learned_with_fallback(net, dmf, b, tol=1e-8)
    assert how in ("learned-fcg", "jacobi-pcg-fallback") and res2.status == 1
    # The preconditioner entry point equals the in-loop one: z = w + CNN(w).
    z = net.precondition(b, diag).cpu().numpy()
    assert np.all(np.isfinite(z)) and np.linalg.norm(z) > 0
