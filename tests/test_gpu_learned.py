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
    """Loss and tied weight gradients of the symmetric network (bf16
    activations on the GPU, float64 oracle): same direction, close size."""
    from cutcell.learned import MaskedCNN

    rng = np.random.default_rng(5)
    nz, ny, nx = 6, 7, 18
    mask = (rng.random((nz, ny, nx)) > 0.2)
    red2full = np.nonzero(mask.ravel())[0].astype(np.int64)
    net = MaskedCNN((nx, ny, nz), red2full, seed=1)
    w0 = 0.5 * rng.standard_normal((27, 16, 16))
    w0[:, 1:, :] = 0.0
    w1 = 0.3 * rng.standard_normal((27, 16, 16))
    net.set_masters(w0, w1)
    layers = net.layer_weights()
    W0, W1 = layers[0], layers[1]
    np.testing.assert_array_equal(layers[2], cnn_ref.flip(W1))
    np.testing.assert_array_equal(layers[3], cnn_ref.flip(W0))
    r = rng.standard_normal(red2full.size)
    e = rng.standard_normal(red2full.size)
    d = 0.5 + rng.random(red2full.size)
    grid = lambda v: _scatter(v, red2full, mask.shape)
    L_ref, g_ref = cnn_ref.sym_loss_and_grads(grid(r), grid(e), grid(d), W0, W1, mask.astype(float))
    diag = _dev(d)
    r32, e32 = r.astype(np.float32), e.astype(np.float32)
    L = float(net.sample_loss_grad(_dev(r32), _dev(e32), diag, 1.0, 1.0, False,
                                   float(np.sum(e32.astype(np.float64) ** 2))).item())
    assert L == pytest.approx(L_ref, rel=2e-2)
    got = net.grad.cpu().numpy().reshape(2, 27, 16, 16)
    for l in range(2):
        a, b = got[l].ravel(), g_ref[l].ravel()
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert cos > 0.99, (l, cos)
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(b), l


def _scatter(v, red2full, shape):
    g = np.zeros(int(np.prod(shape)))
    g[red2full] = v
    return g.reshape(shape)


def test_preconditioner_matches_oracle_and_is_symmetric(lib):
    """z = D^-1 r + S C^T C S r from the HIP entry point against the float64
    oracle (bf16 tolerance), and <M a, b> = <a, M b> to that tolerance."""
    from cutcell.learned import MaskedCNN

    rng = np.random.default_rng(7)
    nz, ny, nx = 7, 9, 20
    mask = rng.random((nz, ny, nx)) > 0.15
    red2full = np.nonzero(mask.ravel())[0].astype(np.int64)
    for periodic in ((False, False, False), (True, False, True)):
        net = MaskedCNN((nx, ny, nz), red2full, periodic=periodic, seed=2)
        w0 = 0.5 * rng.standard_normal((27, 16, 16))
        w0[:, 1:, :] = 0.0
        net.set_masters(w0, 0.3 * rng.standard_normal((27, 16, 16)))
        W0, W1 = net.layer_weights()[:2]
        d = 0.5 + rng.random(red2full.size)
        diag = _dev(d)
        a = rng.standard_normal(red2full.size)
        b = rng.standard_normal(red2full.size)
        Ma = net.precondition(_dev(a), diag).cpu().numpy()
        Mb = net.precondition(_dev(b), diag).cpu().numpy()
        ref, _ = cnn_ref.sym_network(_scatter(a, red2full, mask.shape), _scatter(d, red2full, mask.shape), W0, W1,
                                     mask.astype(float), periodic)
        ref = ref.ravel()[red2full]
        assert np.linalg.norm(Ma - ref) <= 2e-2 * np.linalg.norm(ref)
        assert abs(Ma @ b - a @ Mb) <= 2e-2 * np.linalg.norm(Ma) * np.linalg.norm(b)
        assert Ma @ a >= np.sum(a * a / d) * (1 - 2e-2)


def test_learned_fcg_trains_and_solves(lib):
    """Pairs from the solver's own iterates, Adam on the GPU (loss falls),
    flexible CG to 1e-8 with the float64 true residual as the judge, and the
    gain: the trained preconditioner needs fewer iterations than Jacobi-PCG
    on the same cut system (a float64 simulation of this recipe goes from
    236 to about 100)."""
    import scipy.sparse.linalg as spla

    from cutcell.cut3d import SphereObstacle
    from cutcell.grid3d import Axis, Grid3D
    from cutcell.learned import MaskedCNN, generate_pairs, solve_learned_with_fallback
    from cutcell.matfree import DeviceMatrixFree, assemble_reduced_csr, build_matrix_free

    grid = Grid3D(tuple(Axis(12, 3) for _ in range(3)))
    mf = build_matrix_free(grid, SphereObstacle.random(4))
    dmf = DeviceMatrixFree(mf)
    rng = np.random.default_rng(0)
    rhs = [_dev(rng.standard_normal(mf.n)) for _ in range(24)]
    R, E = generate_pairs(dmf, rhs, snapshots=[3, 10, 30, 60, 120])
    assert R.shape == (120, mf.n) and torch.isfinite(R).all() and torch.isfinite(E).all()
    net = MaskedCNN(grid.nb, mf.interior, seed=0)
    diag = _dev(mf.diag)
    rep = net.train(R, E, diag, steps=300, batch=8, lr=1e-3)
    print(f"loss {rep.loss_first:.3f} -> {rep.loss_last:.3f} in {rep.seconds:.1f} s", flush=True)
    assert np.mean(rep.history[-20:]) < 0.8 * np.mean(rep.history[:5])
    b = _dev(np.random.default_rng(99).standard_normal(mf.n))
    x, res, hist = net.fcg(dmf, b, tol=1e-8, maxiter=2000, history=True)
    xj, resj = dmf.pcg(b, tol=1e-8, maxiter=2000)
    print(f"learned FCG {res.iterations} iterations, Jacobi-PCG {resj.iterations}", flush=True)
    A = assemble_reduced_csr(mf)
    bh = b.cpu().numpy()
    xs = spla.splu(A.tocsc()).solve(bh)
    assert res.status == 1 and res.true_rel <= 1e-7
    assert res.iterations <= 0.8 * resj.iterations
    xh = x.cpu().numpy()
    assert np.linalg.norm(A @ xh - bh) / np.linalg.norm(bh) == pytest.approx(res.true_rel, rel=1e-5)
    assert np.linalg.norm(xh - xs) / np.linalg.norm(xs) < 1e-5
    x2, res2, how = solve_learned_with_fallback(net, dmf, b, tol=1e-8)
    assert how == "learned-fcg" and res2.status == 1
