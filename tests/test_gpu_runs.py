"""GPU checks of the run encoding (``cc_runs_*``) and the single-vector
product over it (``cc_spmv_runs_*``).

The encoding is integer work and is held bit-exact: the device runs equal a
host encoder written from the definition, and decoding them gives back every
CSR column index. The product is held to the CSR product and to SciPy at
float64 rounding (1e-13 of |A| |x|), float32 at 1e-5.
"""

import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

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


def host_runs(A, max_len=255):
    """Runs of every row from the definition: a run ends where the next
    column is not the previous plus one or the run holds max_len entries."""
    A = A.tocsr()
    rptr, recs = [0], []
    for r in range(A.shape[0]):
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
        i = 0
        while i < cols.size:
            j = i + 1
            while j < cols.size and cols[j] == cols[j - 1] + 1 and j - i < max_len:
                j += 1
            recs.append((int(cols[i]), (i << 8) | (j - i)))
            i = j
        rptr.append(len(recs))
    return np.array(rptr, dtype=np.int64), np.array(recs, dtype=np.uint32).reshape(-1, 2)


def decode(rptr, recs, indptr):
    out = np.empty(int(indptr[-1]), dtype=np.int64)
    for r in range(rptr.size - 1):
        for k in range(rptr[r], rptr[r + 1]):
            start, packed = int(recs[k, 0]), int(recs[k, 1])
            off, ln = packed >> 8, packed & 255
            out[indptr[r] + off: indptr[r] + off + ln] = np.arange(start, start + ln)
    return out


def _planar():
    from cutcell.cut3d import PlanarCut, assemble_levelset_3d, polynomial_preserving_extension_3d, restrict_spd

    cut = PlanarCut.random(0)
    asm = assemble_levelset_3d(cut, 8, 3)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    return A


def _long_runs():
    """Rows with a run of 600 consecutive columns (split at 255), empty rows,
    single entries and runs broken by one missing column."""
    n = 700
    rows, cols = [], []
    rows += [0] * 600
    cols += list(range(50, 650))
    rows += [2, 2, 2, 2]
    cols += [3, 4, 6, 7]
    rows += [5]
    cols += [699]
    rng = np.random.default_rng(3)
    for r in range(10, n, 7):
        c = np.unique(rng.integers(0, n, 40))
        rows += [r] * c.size
        cols += c.tolist()
    A = sp.csr_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.fixture(scope="module")
def mats():
    from cutcell.oblique2d import sliver_system

    return {"planar-p3-N8": _planar(), "sliver-16": sliver_system(16, 4, 1e-2).A.tocsr(), "long-runs": _long_runs()}


@pytest.mark.parametrize("name", ["planar-p3-N8", "sliver-16", "long-runs"])
def test_runs_match_definition_and_decode_to_the_indices(hip, mats, name):
    A = mats[name]
    A.sort_indices()
    Ad = hip.DeviceCSR.from_scipy(A)
    R = hip.DeviceRuns.from_csr(Ad)
    rptr, recs = host_runs(A)
    print(f"{name}: n={A.shape[0]} nnz={A.nnz} runs={R.nruns} ({A.nnz / max(R.nruns, 1):.2f} entries per run)",
          flush=True)
    assert R.nruns == recs.shape[0]
    np.testing.assert_array_equal(R.rptr.cpu().numpy(), rptr)
    got = R.runs.cpu().numpy().view(np.uint32)[: 2 * R.nruns].reshape(-1, 2)
    np.testing.assert_array_equal(got, recs)
    np.testing.assert_array_equal(decode(rptr, got, A.indptr), A.indices)


def test_runs_reject_unsorted_or_repeated_columns(hip):
    A = sp.csr_matrix(np.array([[1.0, 2.0, 0.0], [0.0, 3.0, 4.0], [5.0, 0.0, 6.0]]))
    for bad in ([1, 0], [1, 1]):
        indices = A.indices.copy()
        indices[0], indices[1] = bad
        Ad = hip.DeviceCSR(3, A.nnz, _dev(A.indptr.astype(np.int32)), _dev(indices.astype(np.int32)), _dev(A.data))
        with pytest.raises(hip.HipError):
            hip.DeviceRuns.from_csr(Ad)


@pytest.mark.parametrize("name", ["planar-p3-N8", "sliver-16", "long-runs"])
def test_runs_product_matches_csr_and_scipy(hip, mats, name):
    A = mats[name]
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    R = hip.DeviceRuns.from_csr(Ad)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    scale = max((abs(A) @ np.abs(x)).max(), 1e-300)
    y = hip.spmv_runs(R, _dev(x)).cpu().numpy()
    assert np.abs(y - A @ x).max() <= 1e-13 * scale
    y_csr = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    lib = hip.load_library()
    assert lib.cc_spmv_csr_f64(A.shape[0], hip._ptr(Ad.indptr), hip._ptr(Ad.indices), hip._ptr(Ad.vals),
                               hip._ptr(_dev(x)), hip._ptr(y_csr), hip._stream()) == 0
    assert np.abs(y - y_csr.cpu().numpy()).max() <= 1e-13 * scale
    x32 = x.astype(np.float32)
    y32 = hip.spmv_runs(R, _dev(x32)).cpu().numpy()
    ref32 = A.astype(np.float32).astype(np.float64) @ x32.astype(np.float64)
    assert np.abs(y32 - ref32).max() <= 1e-5 * scale


def test_runs_empty_matrix_is_a_no_op(hip):
    lib = hip.load_library()
    nr = ctypes.c_int64(7)
    assert lib.cc_runs_count(0, None, None, None, ctypes.addressof(nr), hip._stream()) == 0
    assert nr.value == 0
    assert lib.cc_spmv_runs_f64(0, None, None, None, None, None, None, hip._stream()) == 0


@pytest.mark.parametrize("name", ["planar-p3-N8", "sliver-16"])
def test_pcg_with_run_products_matches_oracle(hip, mats, name):
    """Single-right-hand-side Jacobi-PCG with its products from the runs:
    the residual trajectory against the oracle's restatement of the reference
    GPU loop (the parity bar of the CSR solves), the solution within 1e-6, and
    the mixed solve to the float64 true residual."""
    import oracle
    from parity import check_trajectory

    A = mats[name].tocsr()
    b = np.random.default_rng(21).standard_normal(A.shape[0])
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    R = hip.DeviceRuns.from_csr(Ad)
    xd, res, hist = hip.pcg(Ad, _dev(b), tol=1e-8, maxiter=5000, history=True, runs=R)
    r = res[0]
    x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(A, b, tol=1e-8, history=True)
    print(f"{name}: runs PCG {r.iterations} it, oracle {it_or}", flush=True)
    assert r.status == hip.CC_CONVERGED == st_or
    check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
    x = xd.cpu().numpy()
    assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < 1e-6
    xm, resm = hip.pcg(Ad, _dev(b), tol=1e-8, maxiter=5000, precision="mixed", runs=R)
    xm = xm.cpu().numpy()
    assert resm[0].status == hip.CC_CONVERGED
    assert np.linalg.norm(A @ xm - b) / np.linalg.norm(b) <= 1e-7


def test_solve_gpu_cg_takes_run_products(hip, mats):
    """The drop-in solve_gpu_cg (reference solve_gpu_cg semantics) on a
    canonical CSR matrix: converged report, true residual under the
    reference's success bar, solution within 1e-6 of SuperLU."""
    import scipy.sparse.linalg as spla

    from fast_solver_backends import solve_gpu_cg

    A = mats["planar-p3-N8"].tocsr()
    b = np.random.default_rng(22).standard_normal(A.shape[0])
    x, rep = solve_gpu_cg(A, b, tol=1e-10)
    assert rep.success and rep.backend == "gpu" and rep.relative_residual <= 1e-9
    x_lu = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 1e-6
