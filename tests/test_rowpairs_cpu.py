"""CPU checks of the 4-row group pairing of the paired row-block form."""

import numpy as np
import scipy.sparse as sp

from cutcell.rowpairs import group_unions, row_group_pairs


def _planar():
    from cutcell.cut3d import PlanarCut, assemble_levelset_3d, polynomial_preserving_extension_3d, restrict_spd

    asm = assemble_levelset_3d(PlanarCut.random(2), 8, 3)
    agg = polynomial_preserving_extension_3d(asm)
    A = (restrict_spd(asm.K, agg.E) + restrict_spd(asm.H, agg.E)).tocsr()
    A.sort_indices()
    return A


def test_pairs_partition_the_groups_and_shrink_the_union():
    A = _planar()
    U = group_unions(A)
    pairs = row_group_pairs(A)
    ng = len(U)
    seen = np.concatenate([pairs[:, 0], pairs[pairs[:, 1] >= 0, 1]])
    assert np.array_equal(np.sort(seen), np.arange(ng))
    assert np.all(np.diff(pairs[:, 0]) > 0)
    for a, b in pairs:
        if b >= 0:
            assert b > a
            assert np.intersect1d(U[a], U[b]).size > 0.5 * U[a].size
    nu4 = sum(u.size for u in U)
    nu8 = sum(np.union1d(U[a], U[b]).size if b >= 0 else U[a].size for a, b in pairs)
    paired = int(np.count_nonzero(pairs[:, 1] >= 0))
    print(f"planar N=8 p=3: {ng} groups, {paired} pairs, union {nu4} -> {nu8} ({nu4 / A.nnz:.3f} -> {nu8 / A.nnz:.3f} "
          f"of nnz)")
    assert paired > 0.6 * ng / 2
    assert nu8 < 0.85 * nu4


def test_pairs_of_tiny_and_diagonal_matrices():
    assert row_group_pairs(sp.identity(3, format="csr")).tolist() == [[0, -1]]
    # A diagonal matrix: no two groups share a column, so nothing pairs.
    pairs = row_group_pairs(sp.identity(40, format="csr"))
    assert pairs.shape == (10, 2) and np.all(pairs[:, 1] == -1)
