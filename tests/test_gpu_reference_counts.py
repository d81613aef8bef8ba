"""The HIP Jacobi-PCG held to the reference's own published numbers, the
headline path (row-block products, 16 right-hand sides, float64 and mixed
precision) on a 3D B-spline system, and the learned flexible CG on a periodic
bump channel.

Reference numbers (committed as data under tests/golden/):
- ``benchmarks/results/gpu_solver_benchmark_3d.csv``: the bell systems
  N = 20, 28, 36, 44 (p = 2, eps = 1e-3; ``benchmarks/gpu_pcg_benchmark_3d.py:49-57``),
  their unknowns and nnz and the reference GPU loop's iterations
  104 / 156 / 177 / 278 at relative residual 1e-8;
- ``code/results/fast_solver_stress.csv``: 25 sliver systems (p = 4,
  N = 8..32, eps = 1e-5..1e-1) and the iterations of the reference's SciPy CG.

Bars (``tests/parity.py``):
- same system, HIP against the oracle's restatement of the reference GPU
  loop: the residual trajectory check of ``parity.check_trajectory``;
- against the reference's published counts: ``REFERENCE_ITER_GAP``. The
  stopping iteration of CG on these systems moves with rounding alone: the
  FP64 oracle run on a row/column permutation of one stress system stops up
  to 3 iterations apart, and on the sliver systems (E pattern and nnz
  bit-identical to the reference's) the oracle's count differs from the
  reference's CSV by up to 3. The bell systems' E rule breaks distance ties
  by NumPy's unstable argsort, whose order depends on the host's SIMD
  dispatch: no dispatch available here (AVX-512, AVX2, baseline) reproduces
  the published nnz, so the rebuilt bell systems (host-independent stable
  tie order) match the published unknowns exactly and the nnz to 0.1%.
Every gap is printed (run with ``-s`` to see the table).
"""

import csv
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from parity import REFERENCE_ITER_GAP, check_trajectory

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
BELL = list(csv.DictReader((GOLDEN / "reference_gpu_solver_benchmark_3d.csv").open()))
STRESS = list(csv.DictReader((GOLDEN / "reference_fast_solver_stress.csv").open()))


@pytest.fixture(scope="module")
def hip():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip as h

    h.load_library()
    return h


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_stress_rows_against_reference_counts(hip):
    """All 25 sliver systems of the reference's stress sweep through the HIP
    solver: integer sizes exact, iterations within REFERENCE_ITER_GAP of the
    reference's count and on the oracle's trajectory, true residual below
    the reference's success bar."""
    from cutcell.oblique2d import sliver_system

    gaps = []
    for row in STRESS:
        N, eps = int(row["n_elements"]), float(row["epsilon"])
        s = sliver_system(N, 4, eps)
        assert (len(s.active), s.A.shape[0], s.A.nnz) == (int(row["n_active"]), int(row["n_reduced"]), int(row["nnz"]))
        xd, res, hist = hip.pcg(hip.DeviceCSR.from_scipy(s.A), _dev(s.b), tol=1e-8, maxiter=5000, history=True)
        r = res[0]
        _, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=1e-8, history=True)
        ref_it = int(row["iterations"])
        gaps.append((N, eps, ref_it, it_or, r.iterations))
        print(f"stress N={N} eps={eps:g}: reference {ref_it}, oracle {it_or}, HIP {r.iterations} "
              f"(gap {r.iterations - ref_it:+d}), true rel {r.true_rel:.2e}", flush=True)
        assert r.status == hip.CC_CONVERGED == st_or
        check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
        assert r.true_rel <= 1e-7  # the reference's success bar max(10 tol, 1e-14)
        assert abs(r.iterations - ref_it) <= REFERENCE_ITER_GAP, (N, eps, ref_it, r.iterations)
    g = np.array([h - ref for _, _, ref, _, h in gaps])
    print(f"stress gaps HIP - reference: {np.bincount(np.abs(g)).tolist()} (count of |gap| = 0, 1, ...), "
          f"within one: {int((np.abs(g) <= 1).sum())}/{len(g)}", flush=True)


@pytest.mark.parametrize("row", BELL, ids=[f"bell-N{r['n_elements']}" for r in BELL])
def test_bell_published_systems(hip, row):
    """The reference's published 3D benchmark systems, one right-hand side
    b = E^T load, through the HIP solver: unknowns exact, nnz within 0.2%,
    iterations within REFERENCE_ITER_GAP of the reference GPU loop's count
    and on the oracle's trajectory for the same system, residual below 1e-8
    as the reference's, and the solution within 1e-6 of SuperLU's."""
    from cutcell.bell3d import bell_system

    N = int(row["n_elements"])
    s = bell_system(N)
    assert s.A.shape[0] == int(row["n_unknowns"])
    assert abs(s.A.nnz - int(row["nnz"])) / int(row["nnz"]) < 2e-3
    xd, res, hist = hip.pcg(hip.DeviceCSR.from_scipy(s.A), _dev(s.b), tol=1e-8, maxiter=20000, history=True)
    r = res[0]
    x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=1e-8, maxiter=20000, history=True)
    ref_it = int(row["gpu_cg_iterations"])
    print(f"bell N={N}: n={s.A.shape[0]} nnz={s.A.nnz} (published {row['nnz']}); reference {ref_it}, "
          f"oracle {it_or}, HIP {r.iterations} (gap {r.iterations - ref_it:+d}); "
          f"HIP recursive rel {r.recursive_rel:.3e} (published {float(row['gpu_cg_residual']):.3e})", flush=True)
    assert r.status == hip.CC_CONVERGED == st_or
    check_trajectory(hist[:, 0], r.iterations, h_or, it_or, 1e-8)
    assert abs(r.iterations - ref_it) <= REFERENCE_ITER_GAP
    assert r.recursive_rel <= 1e-8 and r.true_rel <= 1e-7
    x = xd.cpu().numpy()
    assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < 1e-6
    if N <= 28:  # SuperLU factor time grows steeply (the reference took 279 s at N = 44)
        xs = spla.splu(s.A.tocsc()).solve(s.b)
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6


@pytest.fixture(scope="module")
def planar3d():
    from cutcell.cut3d import planar_cut_system

    s = planar_cut_system(seed=0, degree=3, n_elements=12)
    B = np.random.default_rng(21).standard_normal((s.A.shape[0], 16))
    return s, B, spla.splu(s.A.tocsc())


def test_headline_path_fp64_row_blocks_3d(hip, planar3d):
    """The headline solve's float64 arm on a 3D planar p = 3 system
    (n = 2814, ~295 nonzeros per row: the launch shapes of the 81k system):
    row-block products, 16 right-hand sides, every column on the oracle's
    trajectory, solutions within 1e-6 of the oracle's and of SuperLU's
    (residual 1e-8 on this kappa ~ 1e4 system)."""
    s, B, lu = planar3d
    Ad = hip.DeviceCSR.from_scipy(s.A, with_f32=True)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    X, res, hist = hip.pcg(Ad, _dev(B), tol=1e-8, maxiter=5000, history=True, rb=rb)
    X = X.cpu().numpy()
    for c in range(16):
        x_or, it_or, st_or, _, h_or = oracle.pcg_gpu_loop(s.A, B[:, c], tol=1e-8, history=True)
        assert res[c].status == hip.CC_CONVERGED == st_or
        check_trajectory(hist[:, c], res[c].iterations, h_or, it_or, 1e-8)
        assert np.linalg.norm(X[:, c] - x_or) / np.linalg.norm(x_or) < 1e-6
        xs = lu.solve(B[:, c])
        assert np.linalg.norm(X[:, c] - xs) / np.linalg.norm(xs) < 1e-6
        assert res[c].true_rel <= 1e-8 * 1.0001


def test_headline_path_mixed_row_blocks_3d(hip, planar3d):
    """The headline solve itself (float32 row-block products inside float64
    iterative refinement, 16 right-hand sides) on the 3D planar p = 3 system:
    every column reaches the float64 true residual 1e-8 (recomputed here on
    the host), refinement sweeps within one of the oracle's mixed scheme and
    inner iterations within 10%, solutions within 1e-6 of SuperLU's."""
    s, B, lu = planar3d
    Ad = hip.DeviceCSR.from_scipy(s.A, with_f32=True)
    rb = hip.DeviceRowBlocks.from_csr(Ad)
    X, res = hip.pcg(Ad, _dev(B), tol=1e-8, precision="mixed", inner_tol=1e-4, rb=rb)
    X = X.cpu().numpy()
    for c in range(16):
        rel = np.linalg.norm(s.A @ X[:, c] - B[:, c]) / np.linalg.norm(B[:, c])
        assert res[c].status == hip.CC_CONVERGED and rel <= 1e-8
        assert res[c].true_rel == pytest.approx(rel, rel=1e-5)
        xs = lu.solve(B[:, c])
        assert np.linalg.norm(X[:, c] - xs) / np.linalg.norm(xs) < 1e-6
        if c < 4:
            _, total, sweeps, st, _ = oracle.mixed_refinement_py(s.A, B[:, c], tol=1e-8, inner_tol=1e-4)
            assert st == oracle.CONVERGED
            assert abs(res[c].outer - sweeps) <= 1
            assert abs(res[c].inner_total - total) <= 0.1 * total + 3


def test_learned_fcg_on_periodic_bump_channel(hip):
    """The pressure operator E^T K E of a small bump channel (periodic in x
    and z, as the 128x64x64 benchmark channel): the multilevel network trained
    on the solver's own pairs, flexible CG against Jacobi-PCG. Zero-sum
    right-hand sides (the operator's null space is the constants): FCG reaches
    the float64 true residual 1e-8 in at most 0.6 of Jacobi-PCG's iterations,
    and its solution equals a tight Jacobi-PCG solution up to a constant; the
    screened operator (mass term 1) is checked against SuperLU."""
    from cutcell.grid3d import Axis, Grid3D
    from cutcell.learned import MultilevelCNN, generate_pairs
    from cutcell.matfree import DeviceMatrixFree, assemble_reduced_csr, build_matrix_free
    from cutcell.workloads import bump_levelset

    grid = Grid3D((Axis(16, 3, 9.0, True), Axis(9, 3, 3.036, False), Axis(8, 3, 4.5, True)))
    periodic = tuple(a.periodic for a in grid.axes)
    rng = np.random.default_rng(7)
    for mass in (0.0, 1.0):
        mf = build_matrix_free(grid, bump_levelset(), mass_coef=mass)
        dmf = DeviceMatrixFree(mf)

        def rhs():
            b = rng.standard_normal(mf.n)
            return _dev(b - b.mean() if mass == 0.0 else b)

        R, E = generate_pairs(dmf, [rhs() for _ in range(24)], snapshots=[3, 10, 30, 60, 120])
        net = MultilevelCNN(grid.nb, mf.interior, periodic=periodic, levels=3, seed=0)
        rep = net.train(R, E, _dev(mf.diag), steps=300, batch=8, lr=2e-3)
        assert np.mean(rep.history[-20:]) < 0.7 * np.mean(rep.history[:5])
        b = rhs()
        x, res = net.fcg(dmf, b, tol=1e-8, maxiter=3000)
        xj, resj = dmf.pcg(b, tol=1e-8, maxiter=3000)
        print(f"bump channel mass={mass}: loss {rep.loss_first:.3f} -> {rep.loss_last:.3f}; "
              f"FCG {res.iterations} it, Jacobi-PCG {resj.iterations} it", flush=True)
        assert res.status == 1 and res.true_rel <= 1e-7
        assert res.iterations <= 0.6 * resj.iterations
        x = x.cpu().numpy()
        bh = b.cpu().numpy()
        if mass == 0.0:
            xt, rest = dmf.pcg(b, tol=1e-13, maxiter=20000)
            assert rest.true_rel <= 1e-11
            xt = xt.cpu().numpy()
            d = (x - x.mean()) - (xt - xt.mean())
            assert np.linalg.norm(d) / np.linalg.norm(xt - xt.mean()) < 1e-5
        else:
            A = assemble_reduced_csr(mf)
            xs = spla.splu(A.tocsc()).solve(bh)
            assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-5
