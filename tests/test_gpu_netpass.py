"""GPU checks of the three-launch network pass (csrc/netpass.hip): the
16x16x32 MFMA lane layout, one level's C^T C on a batch of grids against the
float64 oracle, independence of the launch shape, and the multilevel
preconditioner against the per-level four-layer launches it replaces."""

import ctypes

import numpy as np
import pytest

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


def test_mfma32_lane_layout(lib):
    from cutcell import hip
    from cutcell.learned import from_bf16_bits, to_bf16_bits

    rng = np.random.default_rng(1)
    A = to_bf16_bits(rng.standard_normal((16, 32)))
    B = to_bf16_bits(rng.standard_normal((32, 16)))
    D = torch.zeros(256, dtype=torch.float32, device="cuda")
    Ad, Bd = _dev(A), _dev(B)
    assert lib.cc_mfma32_probe(Ad.data_ptr(), Bd.data_ptr(), D.data_ptr(), hip._stream()) == 0
    ref = from_bf16_bits(A).astype(np.float64) @ from_bf16_bits(B).astype(np.float64)
    np.testing.assert_allclose(D.cpu().numpy().reshape(16, 16), ref, rtol=1e-5, atol=1e-5)


def _level_oracle(u, W0, W1, mask, periodic):
    """Channel 0 of the four layers (W0, W1, flip W1, flip W0) on u."""
    x = np.zeros(u.shape + (16,))
    x[..., 0] = u
    a = x
    for W in cnn_ref.sym_layers(W0, W1):
        a, _ = cnn_ref.conv3(a, W, mask, relu=False, periodic=periodic)
    return a[..., 0]


def _level_net(dims, periodic, mask, rng):
    from cutcell.learned import MaskedCNN

    net = MaskedCNN(dims, None, periodic=periodic, mask=mask.ravel().astype(np.uint8))
    w0 = 0.5 * rng.standard_normal((27, 16, 16))
    w0[:, 1:, :] = 0.0
    net.set_masters(w0, 0.3 * rng.standard_normal((27, 16, 16)))
    return net


def _run_level(lib, net, u):
    from cutcell import hip

    nb = u.shape[0]
    d = net.desc()
    wsb = int(lib.cc_net_level_workspace_bytes(ctypes.addressof(d), nb))
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    ud = _dev(u.astype(np.float32))
    y = torch.full((u.size,), np.nan, dtype=torch.float32, device="cuda")
    st = lib.cc_net_level_forward(ctypes.addressof(d), nb, ud.data_ptr(), y.data_ptr(), ws.data_ptr(), wsb,
                                  hip._stream())
    torch.cuda.synchronize()
    assert st == 0
    return y.cpu().numpy().reshape(u.shape)


LEVEL_GRIDS = [
    ((19, 11, 9), (False, False, False)),  # partial tiles on every axis
    ((16, 10, 12), (True, False, True)),  # periodic x and z
    ((5, 3, 4), (False, True, False)),  # smaller than one tile, periodic y of 3
    ((40, 17, 20), (False, False, False)),  # several tiles per axis
]


@pytest.mark.parametrize("dims,periodic", LEVEL_GRIDS)
def test_level_pass_matches_oracle(lib, dims, periodic):
    """y = C^T C u (channel 0, masked) on 3 grids of one level against the
    float64 oracle on the same bf16-rounded input and weights; the GPU rounds
    the three hidden activations to bf16, so the bar is 1e-2 in norm."""
    from cutcell.learned import from_bf16_bits, to_bf16_bits

    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    mask = rng.random((nz, ny, nx)) > 0.2
    net = _level_net(dims, periodic, mask, rng)
    W0, W1 = net.layer_weights()[:2]
    u = from_bf16_bits(to_bf16_bits(rng.standard_normal((3, nz, ny, nx)))).astype(np.float64)
    got = _run_level(lib, net, u)
    assert np.isfinite(got).all()
    for b in range(3):
        ref = _level_oracle(u[b], W0, W1, mask.astype(float), periodic)
        assert not got[b][~mask].any()
        err = np.linalg.norm(got[b] - ref) / np.linalg.norm(ref)
        print(f"dims {dims} periodic {periodic} grid {b}: rel err {err:.2e}", flush=True)
        assert err <= 1e-2
        assert np.abs(got[b] - ref).max() <= 3e-2 * np.abs(ref).max()


def test_level_pass_does_not_depend_on_launch_shape(lib):
    """Every work item is computed the same way whatever block runs it: one
    item per block, a few blocks with many items each (the next item's input
    prefetched across grid boundaries), and the default give equal bits."""
    rng = np.random.default_rng(12)
    dims, periodic = (37, 20, 18), (False, False, True)
    nx, ny, nz = dims
    mask = rng.random((nz, ny, nx)) > 0.1
    net = _level_net(dims, periodic, mask, rng)
    u = rng.standard_normal((4, nz, ny, nx))
    outs = []
    try:
        for blocks in (0, 1 << 20, 7):
            assert lib.cc_net_set_config(blocks, 0) == 0
            outs.append(_run_level(lib, net, u))
    finally:
        lib.cc_net_set_config(0, 0)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("dims,periodic", [((21, 18, 14), (False, False, False)), ((24, 13, 16), (True, False, True))])
def test_multilevel_pass_matches_per_level_layers(lib, dims, periodic):
    """The multilevel preconditioner through the three-launch pass against
    the same preconditioner through the per-level four-layer launches (both
    bf16 with the same rounding points; only fp32 summation order differs)."""
    from cutcell.learned import MultilevelCNN

    rng = np.random.default_rng(13)
    nx, ny, nz = dims
    mask = rng.random((nz, ny, nx)) > 0.15
    red2full = np.nonzero(mask.ravel())[0].astype(np.int64)
    ml = MultilevelCNN(dims, red2full, periodic=periodic, levels=4, net_levels=(0, 1, 2, 3), seed=4)
    for net in ml.nets.values():
        w0 = 0.5 * rng.standard_normal((27, 16, 16))
        w0[:, 1:, :] = 0.0
        net.set_masters(w0, 0.3 * rng.standard_normal((27, 16, 16)))
    diag = _dev(0.5 + rng.random(red2full.size))
    r = _dev(rng.standard_normal(red2full.size))
    try:
        assert lib.cc_net_set_config(0, 1) == 0
        z_old = ml.precondition(r, diag).cpu().numpy()
        assert lib.cc_net_set_config(0, 0) == 0
        z_new = ml.precondition(r, diag).cpu().numpy()
    finally:
        lib.cc_net_set_config(0, 0)
    err = np.linalg.norm(z_new - z_old) / np.linalg.norm(z_old)
    print(f"multilevel dims {dims}: three-launch vs per-level layers rel diff {err:.2e}", flush=True)
    assert err <= 2e-3
