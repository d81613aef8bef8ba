"""GPU parity of the row-block form (``cc_rb_*``) and of the CG solves that
take their products from it.

The form is checked entry by entry against a host construction written from
its definition (integer maps bit-exact, float64 values exact, float32 values
the rounded float64 ones); its products against the oracle SpMV (1e-13
relative in float64, 1e-5 in float32) for every launch shape the autotuner
may pick; the solves against the oracle's restatement of the reference GPU
loop with the same bars as the CSR solves.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from cutcell.oblique2d import oblique_system, sliver_system
from parity import check_trajectory

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

R = 4


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip as h

    h.load_library()
    return h


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _ragged(n, seed, max_len=200):
    """Rows of length 0 to max_len (empty rows included), random columns."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len, n)
    lens[::37] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, rows.size)
    A = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    return A


@pytest.fixture(scope="module")
def mats():
    return {
        "sliver-16": sliver_system(16, 4, 1e-2).A.tocsr(),
        "C1-oblique-5deg": oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32).A.tocsr(),
        "ragged-701": _ragged(701, 5),  # n not a multiple of 4
        "ragged-6": _ragged(6, 7, 6),
    }


def host_row_blocks(A):
    """The row-block form from its definition: per group of R rows, the
    ascending union of their columns and the dense R-row values there."""
    A = A.tocsr()
    n = A.shape[0]
    ng = -(-n // R)
    gptr = [0]
    cols, vals = [], []
    for g in range(ng):
        rows = range(g * R, min(n, (g + 1) * R))
        u = np.unique(np.concatenate([A.indices[A.indptr[r]:A.indptr[r + 1]] for r in rows] + [np.zeros(0, int)]))
        block = np.zeros((u.size, R))
        for i, r in enumerate(rows):
            seg = slice(A.indptr[r], A.indptr[r + 1])
            block[np.searchsorted(u, A.indices[seg]), i] = A.data[seg]
        cols.append(u)
        vals.append(block)
        gptr.append(gptr[-1] + u.size)
    return np.array(gptr), np.concatenate(cols), np.concatenate(vals).reshape(-1)


@pytest.mark.parametrize("name", ["sliver-16", "C1-oblique-5deg", "ragged-701", "ragged-6"])
def test_row_block_form_matches_definition(hip, mats, name):
    A = mats[name]
    Ad = hip.DeviceCSR.from_scipy(A)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    gptr, gcol, gval = host_row_blocks(A)
    assert rb.ngroups == -(-A.shape[0] // R) and rb.nu == gcol.size
    np.testing.assert_array_equal(rb.gptr.cpu().numpy(), gptr)
    np.testing.assert_array_equal(rb.gcol.cpu().numpy()[: rb.nu], gcol)
    np.testing.assert_array_equal(rb.gval.cpu().numpy()[: rb.nu * R], gval)
    np.testing.assert_array_equal(rb.gval32.cpu().numpy()[: rb.nu * R], gval.astype(np.float32))


def test_row_block_build_rejects_descending_indices(hip):
    A = sp.csr_matrix(np.array([[1.0, 2.0, 0.0], [0.0, 3.0, 4.0], [5.0, 0.0, 6.0]]))
    indices = A.indices.copy()
    indices[0], indices[1] = indices[1], indices[0]  # row 0 now reads columns 1, 0
    Ad = hip.DeviceCSR(3, A.nnz, _dev(A.indptr.astype(np.int32)), _dev(indices.astype(np.int32)), _dev(A.data))
    with pytest.raises(hip.HipError):
        hip.DeviceRowBlocks.from_csr(Ad)


@pytest.mark.parametrize("k,vb", [(1, 8), (4, 8), (16, 8), (2, 4), (16, 4)])
def test_row_block_product_every_launch_shape(hip, mats, k, vb):
    from cutcell.autotune import rb_candidates

    lib = hip.load_library()
    try:
        for name, A in mats.items():
            Ad = hip.DeviceCSR.from_scipy(A)
            rb = hip.DeviceRowBlocks.from_csr(Ad)
            X = np.random.default_rng(k).standard_normal((A.shape[0], k))
            if vb == 4:
                X = X.astype(np.float32)
            A_ref = A.astype(np.float32).astype(np.float64) if vb == 4 else A
            ref = np.stack([oracle.spmv(A_ref.tocsr(), X[:, c].astype(np.float64)) for c in range(k)], axis=1)
            scale = max((np.abs(A_ref) @ np.abs(X.astype(np.float64))).max(), 1e-300)
            tol = 1e-5 if vb == 4 else 1e-13
            seen = 0
            for cand in rb_candidates(k, vb):
                assert lib.cc_rb_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"]) == 0
                Y = hip.spmm_rb(rb, _dev(X)).cpu().numpy().reshape(A.shape[0], k)
                assert np.abs(Y - ref).max() <= tol * scale, (name, cand)
                seen += 1
            assert seen >= 4
    finally:
        lib.cc_rb_set_config(k, vb, 0, 0, 0, 0)


@pytest.mark.parametrize("name", ["sliver-16", "C1-oblique-5deg"])
def test_pcg_row_blocks_matches_oracle(hip, mats, name):
    s = {"sliver-16": sliver_system(16, 4, 1e-2), "C1-oblique-5deg": oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32)}[name]
    Ad = hip.DeviceCSR.from_scipy(s.A)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    xd, res, hist = hip.pcg(Ad, _dev(s.b), tol=1e-8, maxiter=5000, history=True, rb=rb)
    r = res[0]
    x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=1e-8, history=True)
    assert r.status == hip.CC_CONVERGED == st_or
    check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
    x = xd.cpu().numpy()
    assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < 1e-6
    assert r.true_rel == pytest.approx(np.linalg.norm(s.A @ x - s.b) / np.linalg.norm(s.b), rel=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_row_blocks_multi_rhs(hip, precision):
    """16 right-hand sides: every column reaches the float64 true residual and
    agrees with the CSR solve of the same precision."""
    s = oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32)
    Ad = hip.DeviceCSR.from_scipy(s.A, with_f32=True)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    B = np.random.default_rng(11).standard_normal((s.A.shape[0], 16))
    X_rb, res_rb = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision, rb=rb)
    X_csr, res_csr = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision)
    X_rb = X_rb.cpu().numpy()
    X_csr = X_csr.cpu().numpy()
    for c in range(16):
        assert res_rb[c].status == hip.CC_CONVERGED
        rel = np.linalg.norm(s.A @ X_rb[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
        assert rel <= 1e-7
        assert res_rb[c].true_rel == pytest.approx(rel, rel=1e-5, abs=1e-12)
        # Summation order differs from the CSR product's: iterations within
        # the trajectory bar (per refinement sweep for mixed), solutions to
        # the accuracy a 1e-8 residual leaves.
        assert abs(res_rb[c].iterations - res_csr[c].iterations) <= (3 if precision == "fp64" else 6)
        assert np.linalg.norm(X_rb[:, c] - X_csr[:, c]) / np.linalg.norm(X_csr[:, c]) < (
            1e-6 if precision == "fp64" else 1e-4)


def host_union_runs(gptr, gcol):
    """Runs of every group's union columns, from the definition (a run ends
    where the next column is not the previous plus one, or at 255 entries)."""
    rptr, recs = [0], []
    for g in range(gptr.size - 1):
        cols = gcol[gptr[g]:gptr[g + 1]]
        i = 0
        while i < cols.size:
            j = i + 1
            while j < cols.size and cols[j] == cols[j - 1] + 1 and j - i < 255:
                j += 1
            recs.append((int(cols[i]), (i << 8) | (j - i)))
            i = j
        rptr.append(len(recs))
    return np.array(rptr), np.array(recs, dtype=np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("k,vb", [(1, 8), (16, 8), (4, 4), (16, 4)])
def test_row_block_product_from_runs(hip, mats, k, vb):
    """With the union columns given as runs, the one-wave-per-group shapes
    read them from the runs: runs bit-exact against the definition, products
    against the oracle at the gcol product's bar for every launch shape (the
    other shapes keep reading gcol)."""
    from cutcell.autotune import rb_candidates

    lib = hip.load_library()
    try:
        for name, A in mats.items():
            Ad = hip.DeviceCSR.from_scipy(A)
            rb = hip.DeviceRowBlocks.from_csr(Ad, runs=True)
            gptr, gcol, _ = host_row_blocks(A)
            rptr, recs = host_union_runs(gptr, gcol)
            assert rb.nruns == recs.shape[0]
            np.testing.assert_array_equal(rb.grptr.cpu().numpy(), rptr)
            np.testing.assert_array_equal(rb.grun.cpu().numpy().view(np.uint32)[: 2 * rb.nruns].reshape(-1, 2), recs)
            X = np.random.default_rng(k + 1).standard_normal((A.shape[0], k))
            if vb == 4:
                X = X.astype(np.float32)
            A_ref = A.astype(np.float32).astype(np.float64) if vb == 4 else A
            ref = np.stack([oracle.spmv(A_ref.tocsr(), X[:, c].astype(np.float64)) for c in range(k)], axis=1)
            scale = max((np.abs(A_ref) @ np.abs(X.astype(np.float64))).max(), 1e-300)
            tol = 1e-5 if vb == 4 else 1e-13
            for cand in rb_candidates(k, vb):
                assert lib.cc_rb_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"]) == 0
                Y = hip.spmm_rb(rb, _dev(X)).cpu().numpy().reshape(A.shape[0], k)
                assert np.abs(Y - ref).max() <= tol * scale, (name, cand)
    finally:
        lib.cc_rb_set_config(k, vb, 0, 0, 0, 0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_row_block_runs_on_a_3d_planar_system(hip, precision):
    """The headline's solve with the row-block products read from runs, on a
    3D planar p=3 system, against the same solve reading gcol: same status,
    iterations within the rounding bar, solutions within 1e-6 (fp64) / 1e-4
    (mixed) and every column at the float64 true residual."""
    from cutcell.cut3d import PlanarCut, assemble_levelset_3d, polynomial_preserving_extension_3d, restrict_spd

    asm = assemble_levelset_3d(PlanarCut.random(1), 9, 3)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    B = np.random.default_rng(12).standard_normal((A.shape[0], 16))
    rb_runs = hip.DeviceRowBlocks.from_csr(Ad, runs=True)
    rb_cols = hip.DeviceRowBlocks.from_csr(Ad)
    X_r, res_r = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision, rb=rb_runs)
    X_c, res_c = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision, rb=rb_cols)
    X_r, X_c = X_r.cpu().numpy(), X_c.cpu().numpy()
    print(f"planar N=9 p=3 n={A.shape[0]} nnz={A.nnz} runs={rb_runs.nruns} union={rb_runs.nu}: "
          f"iterations runs {[r.iterations for r in res_r][:4]} gcol {[r.iterations for r in res_c][:4]}", flush=True)
    for c in range(16):
        assert res_r[c].status == hip.CC_CONVERGED
        assert np.linalg.norm(A @ X_r[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-7
        assert abs(res_r[c].iterations - res_c[c].iterations) <= (3 if precision == "fp64" else 6)
        assert np.linalg.norm(X_r[:, c] - X_c[:, c]) / np.linalg.norm(X_c[:, c]) < (1e-6 if precision == "fp64" else 1e-4)


def host_paired_blocks(A, pairs):
    """The paired row-block form from its definition: per pair, the
    ascending union of its (up to) 8 rows' columns and the dense 8-row values
    there (rows of a missing second group, or past n, are zero)."""
    A = A.tocsr()
    n = A.shape[0]
    gptr, cols, vals = [0], [], []
    for a, b in pairs:
        rows = [a * R + i if a * R + i < n else -1 for i in range(R)]
        rows += [b * R + i if b >= 0 and b * R + i < n else -1 for i in range(R)]
        segs = [A.indices[A.indptr[r]:A.indptr[r + 1]] for r in rows if r >= 0]
        u = np.unique(np.concatenate(segs + [np.zeros(0, int)]))
        block = np.zeros((u.size, 2 * R))
        for i, r in enumerate(rows):
            if r >= 0:
                seg = slice(A.indptr[r], A.indptr[r + 1])
                block[np.searchsorted(u, A.indices[seg]), i] = A.data[seg]
        cols.append(u)
        vals.append(block)
        gptr.append(gptr[-1] + u.size)
    return np.array(gptr), np.concatenate(cols), np.concatenate(vals).reshape(-1)


def _planar_csr(seed, N):
    from cutcell.cut3d import PlanarCut, assemble_levelset_3d, polynomial_preserving_extension_3d, restrict_spd

    asm = assemble_levelset_3d(PlanarCut.random(seed), N, 3)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("k,vb", [(1, 8), (16, 8), (4, 4), (16, 4)])
def test_paired_row_blocks_match_definition_and_oracle(hip, mats, k, vb):
    """Paired row blocks: form entry by entry against the definition
    (integers bit-exact, float64 values exact, float32 the rounded ones), the
    product against the oracle for every launch shape."""
    from cutcell.autotune import rb_candidates
    from cutcell.rowpairs import row_group_pairs

    lib = hip.load_library()
    cases = dict(mats)
    cases["planar-p3-N8"] = _planar_csr(2, 8)
    try:
        for name, A in cases.items():
            A = A.tocsr()
            A.sort_indices()
            pairs = row_group_pairs(A)
            Ad = hip.DeviceCSR.from_scipy(A)
            rb = hip.DeviceRowBlocks.from_csr_pairs(Ad, pairs)
            gptr, gcol, gval = host_paired_blocks(A, pairs)
            assert rb.ngroups == pairs.shape[0] and rb.nu == gcol.size
            np.testing.assert_array_equal(rb.gptr.cpu().numpy(), gptr)
            np.testing.assert_array_equal(rb.gcol.cpu().numpy()[: rb.nu], gcol)
            np.testing.assert_array_equal(rb.gval.cpu().numpy()[: rb.nu * 2 * R], gval)
            np.testing.assert_array_equal(rb.gval32.cpu().numpy()[: rb.nu * 2 * R], gval.astype(np.float32))
            X = np.random.default_rng(k + 2).standard_normal((A.shape[0], k))
            if vb == 4:
                X = X.astype(np.float32)
            A_ref = A.astype(np.float32).astype(np.float64) if vb == 4 else A
            ref = np.stack([oracle.spmv(A_ref.tocsr(), X[:, c].astype(np.float64)) for c in range(k)], axis=1)
            scale = max((np.abs(A_ref) @ np.abs(X.astype(np.float64))).max(), 1e-300)
            tol = 1e-5 if vb == 4 else 1e-13
            for cand in rb_candidates(k, vb):
                assert lib.cc_rb_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"]) == 0
                Y = hip.spmm_rb(rb, _dev(X)).cpu().numpy().reshape(A.shape[0], k)
                assert np.abs(Y - ref).max() <= tol * scale, (name, cand)
    finally:
        lib.cc_rb_set_config(k, vb, 0, 0, 0, 0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pcg_paired_row_blocks_on_a_3d_planar_system(hip, precision):
    """The headline solve with paired row blocks on a 3D planar p=3 system
    against the 4-row form: same status, iterations within the rounding bar,
    solutions within 1e-6 (fp64) / 1e-4 (mixed), float64 true residuals."""
    from cutcell.rowpairs import row_group_pairs

    A = _planar_csr(1, 9)
    Ad = hip.DeviceCSR.from_scipy(A, with_f32=True)
    B = np.random.default_rng(13).standard_normal((A.shape[0], 16))
    rb8 = hip.DeviceRowBlocks.from_csr_pairs(Ad, row_group_pairs(A))
    rb4 = hip.DeviceRowBlocks.from_csr(Ad)
    X8, res8 = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision, rb=rb8)
    X4, res4 = hip.pcg(Ad, _dev(B), tol=1e-8, precision=precision, rb=rb4)
    X8, X4 = X8.cpu().numpy(), X4.cpu().numpy()
    print(f"planar N=9 p=3 n={A.shape[0]}: union 4-row {rb4.nu}, paired {rb8.nu}; iterations paired "
          f"{[r.iterations for r in res8][:4]} 4-row {[r.iterations for r in res4][:4]}", flush=True)
    for c in range(16):
        assert res8[c].status == hip.CC_CONVERGED
        assert np.linalg.norm(A @ X8[:, c] - B[:, c]) / np.linalg.norm(B[:, c]) <= 1e-7
        if precision == "fp64":
            assert abs(res8[c].iterations - res4[c].iterations) <= 3
        elif res8[c].outer == res4[c].outer:
            assert abs(res8[c].iterations - res4[c].iterations) <= 6
        else:
            # Another summation order can leave a sweep's float64 residual
            # just above tol (one more sweep) or just below it (one fewer).
            assert abs(res8[c].outer - res4[c].outer) == 1
        assert np.linalg.norm(X8[:, c] - X4[:, c]) / np.linalg.norm(X4[:, c]) < (1e-6 if precision == "fp64" else 1e-4)
