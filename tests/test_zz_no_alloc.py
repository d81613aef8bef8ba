"""cc_pcg_jacobi_f64_noalloc (fixed device scratch, no workspace argument)
against cc_pcg_jacobi_f64 (caller workspace), both called through the C ABI.

The two run the same kernels in the same order, so iterations, statuses,
residuals, the residual histories and the solutions must be bit-identical.
Two calls in a row, the second on a smaller system with fewer right-hand
sides, show that the scratch left by the first call does not leak into the
second; a problem larger than the scratch is rejected before any GPU work.
"""

import ctypes

import numpy as np
import pytest

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


def _solve(hip, A, B, noalloc, maxiter=5000):
    lib = hip.load_library()
    n, k = B.shape
    Ad = hip.DeviceCSR.from_scipy(A)
    Bd = _dev(B)
    X = torch.full((n, k), float("nan"), dtype=torch.float64, device="cuda")
    hist = torch.full((maxiter + 1, k), float("nan"), dtype=torch.float64, device="cuda")
    res = (hip.PcgResult * k)()
    s = int(torch.cuda.current_stream().cuda_stream)
    args_head = (n, k, int(Ad.indptr.data_ptr()), int(Ad.indices.data_ptr()), int(Ad.vals.data_ptr()),
                 int(Bd.data_ptr()), int(X.data_ptr()), 1e-8, maxiter, 1, 16)
    if noalloc:
        st = lib.cc_pcg_jacobi_f64_noalloc(*args_head, ctypes.addressof(res), int(hist.data_ptr()), s)
    else:
        nbytes = int(lib.cc_pcg_workspace_bytes(n, k, 0))
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        st = lib.cc_pcg_jacobi_f64(*args_head, int(ws.data_ptr()), nbytes, ctypes.addressof(res),
                                   int(hist.data_ptr()), s)
    assert st == 0, st
    torch.cuda.synchronize()
    return X.cpu().numpy(), [(r.iterations, r.status, r.recursive_rel, r.true_rel) for r in res], hist.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(hip):
    from cutcell.oblique2d import oblique_system, sliver_system

    big = oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32)
    small = sliver_system(8, 4, 1e-3)
    rng = np.random.default_rng(3)
    cases = [(big.A, rng.standard_normal((big.A.shape[0], 16))),
             (small.A, rng.standard_normal((small.A.shape[0], 4)))]
    for A, B in cases:
        print(f"noalloc: n={A.shape[0]} k={B.shape[1]}", flush=True)
        X0, r0, h0 = _solve(hip, A, B, noalloc=False)
        X1, r1, h1 = _solve(hip, A, B, noalloc=True)
        assert r1 == r0
        assert all(st == hip.CC_CONVERGED for _, st, _, _ in r1)
        np.testing.assert_array_equal(X1, X0)
        np.testing.assert_array_equal(h1, h0)
        for c in range(B.shape[1]):
            rel = np.linalg.norm(A @ X1[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
            assert rel <= 1e-7


def test_noalloc_rejects_problems_beyond_the_scratch(hip):
    lib = hip.load_library()
    cap = int(lib.cc_pcg_noalloc_max_bytes())
    assert cap == 64 << 20
    n = 200_000  # 16 columns need 5 * n * 16 * 8 bytes > 64 MiB
    assert int(lib.cc_pcg_workspace_bytes(n, 16, 0)) > cap
    dummy = torch.zeros(4, dtype=torch.float64, device="cuda")
    res = (hip.PcgResult * 16)()
    p = int(dummy.data_ptr())
    st = lib.cc_pcg_jacobi_f64_noalloc(n, 16, p, p, p, p, p, 1e-8, 10, 1, 16, ctypes.addressof(res), None,
                                       int(torch.cuda.current_stream().cuda_stream))
    assert st == hip.CC_E_TOO_LARGE
    with pytest.raises(hip.HipError, match="scratch"):
        hip._check(st, "pcg_noalloc")
