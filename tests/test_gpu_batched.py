"""GPU parity of the batched persistent solve against the oracle, per system."""

import numpy as np
import pytest

import oracle
from cutcell.workloads import c3_batch
from parity import STOP_SLACK, check_stop_gap

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell.batched import BatchedSystems

    systems = c3_batch(n_elements=12, n_cuts=4, degrees=(3, 4, 5), n_slivers=3, seed=5, jobs=1)
    return systems, BatchedSystems([s.A for s in systems])


def _check_against_oracle(s, x, r, tol, rel_x=1e-6):
    x_or, it_or, st_or, rel_or, h_or = oracle.pcg_gpu_loop(s.A, s.b, tol=tol, history=True)
    assert r.status == 1 == st_or
    check_stop_gap(r.iterations, it_or, system=(s.A, s.b), tol=tol)
    if r.iterations < it_or:
        assert h_or[r.iterations] <= STOP_SLACK * tol
    assert r.recursive_rel <= tol
    assert r.true_rel <= 10 * tol
    host = np.linalg.norm(s.A @ x - s.b) / np.linalg.norm(s.b)
    assert r.true_rel == pytest.approx(host, rel=1e-6)
    assert np.linalg.norm(x - x_or) / np.linalg.norm(x_or) < rel_x


@pytest.mark.parametrize("block,team", [(256, 16), (512, 32), (1024, 8)])
def test_batched_fp64_matches_oracle(batch, block, team):
    systems, bs = batch
    out = bs.solve([s.b for s in systems], tol=1e-8, block_threads=block, team_lanes=team)
    assert len(out.results) == len(systems)
    for s, x, r in zip(systems, out.X, out.results):
        _check_against_oracle(s, x, r, 1e-8)


def test_batched_mixed_reaches_fp64_tolerance(batch):
    import scipy.sparse.linalg as spla

    systems, bs = batch
    out = bs.solve([s.b for s in systems], tol=1e-8, precision="mixed", block_threads=512, team_lanes=16)
    for s, x, r in zip(systems, out.X, out.results):
        rel = np.linalg.norm(s.A @ x - s.b) / np.linalg.norm(s.b)
        assert r.status == 1 and rel <= 1e-8 and r.outer >= 2
        assert r.true_rel == pytest.approx(rel, rel=1e-5)
        xs = spla.splu(s.A.tocsc()).solve(s.b)
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-5


def test_batched_autotuned_and_repeatable(batch, tmp_path, monkeypatch):
    monkeypatch.setenv("CUTCELL_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    from cutcell import autotune

    autotune._MEM.clear()
    systems, bs = batch
    cfg = autotune.batched_config(bs, "fp64", force=True, tune_iters=10)
    assert cfg["block_threads"] in (256, 512, 1024) and cfg["team_lanes"] in (8, 16, 32, 64)
    assert (tmp_path / "tune.json").exists()
    a = bs.solve([s.b for s in systems], tol=1e-8)
    b = bs.solve([s.b for s in systems], tol=1e-8)
    for xa, xb, ra, rb in zip(a.X, b.X, a.results, b.results):
        np.testing.assert_array_equal(xa, xb)  # deterministic reductions
        assert ra.iterations == rb.iterations
    autotune._MEM.clear()


def test_batched_edge_cases():
    import scipy.sparse as sp

    from cutcell.batched import BatchedSystems

    mats = [sp.eye(3, format="csr"), -sp.eye(4, format="csr"), sp.diags([2.0, 3.0], format="csr")]
    bs = BatchedSystems(mats)
    out = bs.solve([np.zeros(3), np.ones(4), np.array([1.0, 1.0])], tol=1e-12, block_threads=256, team_lanes=8)
    r0, r1, r2 = out.results
    assert (r0.iterations, r0.status) == (1, 2)  # zero right-hand side: breakdown at k = 1
    assert (r1.iterations, r1.status) == (1, 2)  # negative definite
    assert (r2.iterations, r2.status) == (1, 1)  # Jacobi-exact diagonal system
    np.testing.assert_allclose(out.X[2], [0.5, 1.0 / 3.0], rtol=1e-15)
    with pytest.raises(ValueError):
        BatchedSystems([sp.eye(9000, format="csr")])


def test_batched_rejects_a_system_beyond_n_max():
    """A direct caller passing n_max below a system's size: the launch shape
    is chosen from n_max (256 threads, one row each), the 300-row system gets
    status CC_E_ARG (-1) and an untouched solution, the others still solve."""
    import scipy.sparse as sp
    import torch

    from cutcell.batched import BatchedSystems

    mats = [sp.diags(np.full(300, 2.0), format="csr"), sp.diags([2.0, 4.0], format="csr")]
    bs = BatchedSystems(mats)
    bs.n_max = 100  # what a careless direct ABI caller might pass
    B = bs.concat([np.ones(300), np.ones(2)])
    X = torch.full_like(B, 7.0)
    bs.solve_device(B, X, tol=1e-12, block_threads=256, team_lanes=8)
    torch.cuda.synchronize()
    r_big, r_small = bs.read_results()
    assert r_big.status == -1 and r_big.iterations == 0
    assert r_small.status == 1
    Xs = bs.split(X)
    assert np.all(Xs[0] == 7.0)
    np.testing.assert_allclose(Xs[1], [0.5, 0.25], rtol=1e-15)
