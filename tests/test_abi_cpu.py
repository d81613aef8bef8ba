"""CPU checks of the C ABI library: it loads and exports every declared symbol."""

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _header_symbols():
    text = (ROOT / "include" / "cutcell_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cutcell import hip

    lib = hip.load_library()
    declared = _header_symbols()
    assert declared, "no entry points parsed from the header"
    for name in declared:
        assert hasattr(lib, name), name
    assert set(hip.declared_symbols()) <= set(declared)
    assert lib.cc_version() >= 10000


def test_workspace_sizes_are_monotone():
    from cutcell import hip

    lib = hip.load_library()
    a = lib.cc_pcg_workspace_bytes(1000, 1, 0)
    b = lib.cc_pcg_workspace_bytes(2000, 1, 0)
    c = lib.cc_pcg_workspace_bytes(1000, 16, 0)
    m = lib.cc_pcg_workspace_bytes(1000, 16, 1)
    assert 0 < a < b and a < c and m > 0
    assert lib.cc_pcg_workspace_bytes(1000, 3, 0) == lib.cc_pcg_workspace_bytes(1000, 3, 0)
    assert lib.cc_pcg_workspace_bytes(1000, 3, 1) == 0  # k must be 1, 2, 4, 8 or 16


def test_bad_arguments_are_rejected_without_a_gpu():
    from cutcell import hip

    lib = hip.load_library()
    assert lib.cc_spmv_csr_f64(-1, None, None, None, None, None, None) == -1
    assert lib.cc_pcg_jacobi_f64(0, 1, None, None, None, None, None, 1e-8, 10, 1, 16, None, 0, None, None, None) == -1


def test_spmm_config_is_validated_per_vector_width():
    """Launch shapes are kept per value width: four float64 right-hand sides
    do not fit one 16-byte load, four float32 ones do."""
    from cutcell import hip

    lib = hip.load_library()
    try:
        assert lib.cc_spmm_set_config(16, 8, 16, 0, 0, 4, 0, 0) == -1
        assert lib.cc_spmm_set_config(16, 4, 16, 0, 0, 4, 0, 0) == 0
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 0, 0) == 0
        assert lib.cc_spmm_set_config(16, 8, 16, 512, 4, 2, 0, 0) == -1  # 8 x 16 lanes > 64
        assert lib.cc_spmm_set_config(3, 8, 0, 0, 0, 0, 0, 0) == -1
        assert lib.cc_spmm_set_config(16, 2, 0, 0, 0, 0, 0, 0) == -1
        assert lib.cc_spmm_set_config(16, 8, 0, 0, 3, 0, 0, 0) == -1
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 1, 0) == 0
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 1, 2, 1, 0) == -1  # nontemporal needs unroll 4
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 2, 0) == -1
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 8, 2, 1, 0) == 0
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 2, 2, 0, 0) == -1
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 0, 64) == 0  # sweep steps 0..64
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 0, 65) == -1
        assert lib.cc_spmm_set_config(16, 8, 8, 512, 4, 2, 0, -1) == -1
        assert lib.cc_spmm_set_config(1, 8, 64, 0, 16, 1, 1, 8) == 0  # window kernel: k = 1 only
        assert lib.cc_spmm_set_config(2, 8, 32, 0, 16, 1, 0, 0) == -1
    finally:
        for k in hip.SUPPORTED_K:
            for vb in (8, 4):
                assert lib.cc_spmm_set_config(k, vb, 0, 0, 0, 0, 0, 0) == 0


def test_gpu_backend_without_gpu_reports_unavailable():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import numpy as np
    import scipy.sparse as sp

    from fast_solver_backends import solve_gpu_cg, solve_spd

    A = sp.diags([4.0, 5.0, 6.0], format="csr")
    x, rep = solve_gpu_cg(A, np.ones(3))
    assert x is None and not rep.success and rep.message.startswith("HIP backend unavailable")
    x, rep = solve_spd(A, np.ones(3), backend="auto", tol=1e-10, maxiter=20)
    assert rep.backend == "cpu" and rep.success and rep.message.startswith("GPU skipped/fell back")


def test_noalloc_entry_rejects_oversize_before_touching_the_gpu():
    """cc_pcg_jacobi_f64_noalloc checks the problem against its fixed 64 MiB
    scratch before any HIP call (the pointers are never dereferenced)."""
    import ctypes

    from cutcell import hip

    lib = hip.load_library()
    assert lib.cc_pcg_noalloc_max_bytes() == 64 << 20
    dummy = ctypes.c_double(0.0)
    p = ctypes.addressof(dummy)
    res = (hip.PcgResult * 16)()
    st = lib.cc_pcg_jacobi_f64_noalloc(200_000, 16, p, p, p, p, p, 1e-8, 10, 1, 16, ctypes.addressof(res), None, None)
    assert st == hip.CC_E_TOO_LARGE
    assert lib.cc_pcg_jacobi_f64_noalloc(10, 3, p, p, p, p, p, 1e-8, 10, 1, 16, ctypes.addressof(res), None, None) == -1


def test_cpu_gmres_follows_scipy():
    """solve_cpu_gmres (code/fast_solver_backends.py:135-191): restarted SciPy
    GMRES with the Jacobi preconditioner, iterations counted by the pr_norm
    callback, the reference's success rule; on a nonsymmetric
    convection-diffusion system its count and solution equal a direct SciPy
    call's, and it reports failure when maxiter is too small."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from fast_solver_backends import solve_cpu_gmres

    n = 300
    A = sp.diags([-1.3 * np.ones(n - 1), 2.5 * np.ones(n), -0.7 * np.ones(n - 1)], [-1, 0, 1], format="csr")
    b = np.random.default_rng(0).standard_normal(n)
    x, rep = solve_cpu_gmres(A, b, tol=1e-10, restart=20)
    count = [0]
    inv = 1.0 / A.diagonal()
    M = spla.LinearOperator(A.shape, matvec=lambda r: inv * r)
    xs, info = spla.gmres(A, b, rtol=1e-10, atol=0.0, restart=20, maxiter=5000, M=M,
                          callback=lambda _: count.__setitem__(0, count[0] + 1), callback_type="pr_norm")
    assert info == 0 and rep.success and rep.method == "scipy.sparse.linalg.gmres"
    assert rep.iterations == count[0]
    np.testing.assert_array_equal(x, xs)
    assert rep.relative_residual <= 1e-9
    _, rep2 = solve_cpu_gmres(A, b, tol=1e-12, restart=5, maxiter=1)
    assert not rep2.success and rep2.message.startswith("not converged")
