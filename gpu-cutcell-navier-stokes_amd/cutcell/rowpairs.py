"""Pairing of 4-row groups for the paired row-block form (``cc_rb_desc.pairs``).

The row-block product gathers one row of X per union column of a group. On
the 3D B-spline systems the group of 4 consecutive rows (consecutive x) that
shares most of a group's columns is the group of the same x-range a few grid
lines away in y; merging the two into one 8-row group cuts the union from
0.36 to about 0.23-0.25 of the nonzeros on the 81k planar system, and the
gathers with it.

The pairing is host work on the sparsity pattern: rank group offsets
``d = 1 .. max_offset`` by their mean column overlap on a sample of groups,
then, offset by offset in that order, pair each unpaired group ``g`` with
``g + d`` when that group is unpaired too and they share more than
``min_overlap`` of ``g``'s columns. Groups left over stay alone.
"""

from __future__ import annotations

import numpy as np

GROUP_ROWS = 4  # CC_RB_ROWS of the C header


def group_unions(A, rows: int = GROUP_ROWS) -> list[np.ndarray]:
    """Ascending union of the columns of each group of ``rows`` rows."""
    A = A.tocsr()
    n = A.shape[0]
    ng = -(-n // rows)
    ip, ix = A.indptr, A.indices
    return [np.unique(ix[ip[g * rows]:ip[min(n, (g + 1) * rows)]]) for g in range(ng)]


def _overlap(a: np.ndarray, b: np.ndarray) -> int:
    return int(np.intersect1d(a, b, assume_unique=True).size)


def row_group_pairs(A, max_offset: int = 64, passes: int = 4, sample: int = 400, min_overlap: float = 0.5,
                    seed: int = 0) -> np.ndarray:
    """int32 ``[npairs, 2]``: the two 4-row groups of each 8-row group (the
    second -1 for a group left alone), every group exactly once, ordered by
    the first group."""
    U = group_unions(A)
    ng = len(U)
    if ng == 1:
        return np.array([[0, -1]], dtype=np.int32)
    rng = np.random.default_rng(seed)
    span = min(max_offset, ng - 1)
    cand = np.arange(ng - span) if ng - span <= sample else rng.choice(ng - span, sample, replace=False)
    score = {}
    for d in range(1, span + 1):
        fr = [_overlap(U[g], U[g + d]) / max(U[g].size, 1) for g in cand]
        score[d] = float(np.mean(fr)) if fr else 0.0
    order = sorted(score, key=lambda d: -score[d])[:passes]
    partner = np.full(ng, -2, dtype=np.int64)  # -2: unpaired
    for d in order:
        for g in range(ng - d):
            h = g + d
            if partner[g] != -2 or partner[h] != -2:
                continue
            if _overlap(U[g], U[h]) > min_overlap * U[g].size:
                partner[g], partner[h] = h, g
    out = []
    for g in range(ng):
        if partner[g] == -2:
            out.append((g, -1))
        elif partner[g] > g:
            out.append((g, int(partner[g])))
    pairs = np.array(out, dtype=np.int32).reshape(-1, 2)
    seen = np.concatenate([pairs[:, 0], pairs[pairs[:, 1] >= 0, 1]])
    assert seen.size == ng and np.array_equal(np.sort(seen), np.arange(ng))
    return pairs
