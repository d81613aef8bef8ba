"""Shared parity criteria for iterative solves.

Iteration counts of CG on these ill-conditioned cut-cell systems move with
the floating-point summation order alone: the oracle run on a row/column
permutation of the same system stops up to three iterations apart, with
residual envelopes within a factor ~1.7 (measured on the sliver systems,
tests/test_oracle.py::test_summation_order_moves_the_stop_by_rounding_only).
The parity check therefore compares the whole residual trajectory and asks
that any difference in the stopping iteration be explained by the residual
hovering at the threshold: at the iteration where one run stopped, the other
run's residual must lie within the same envelope factor of the tolerance
(CG residual norms are not monotone, so a run one iteration short of its stop
can sit well above the threshold; 1.8x was seen for batched solves of the
2D oblique systems).
"""

import numpy as np

# Measured on the GPU suite: the largest stop gap between a HIP solve and the
# oracle (or between two HIP solves of one system that sum in different
# orders) is 3, over 77 trajectories; 33 of them stop on the same iteration.
# Permuting one stress system in the FP64 oracle alone moves its stop by 3.
MAX_ITER_GAP = 3
ENVELOPE_FACTOR = 2.0
# Against the reference's published counts: at most 3 on the 25 stress rows
# (19 within one) and 2 on the bell systems N = 20..44 (rebuilt to 0.1% of
# the published nnz).
REFERENCE_ITER_GAP = 3
STOP_SLACK = ENVELOPE_FACTOR


GAPS: list = []  # every stop gap accepted in this session (printed at the end)


def check_trajectory(h_a, it_a, h_b, it_b, tol):
    """Assert two residual histories (index = iteration) describe the same solve."""
    h_a = np.asarray(h_a, dtype=float)
    h_b = np.asarray(h_b, dtype=float)
    assert abs(it_a - it_b) <= MAX_ITER_GAP, (it_a, it_b)
    m = min(it_a, it_b)
    e_a = np.minimum.accumulate(h_a[: m + 1])
    e_b = np.minimum.accumulate(h_b[: m + 1])
    ratio = np.max(np.abs(np.log(np.maximum(e_a, 1e-300) / np.maximum(e_b, 1e-300))))
    assert ratio <= np.log(ENVELOPE_FACTOR), ratio
    # Whoever stopped first did so where the other was also at the threshold.
    if it_a < it_b:
        assert h_b[it_a] <= STOP_SLACK * tol, (it_a, it_b, h_b[it_a])
    elif it_b < it_a:
        assert h_a[it_b] <= STOP_SLACK * tol, (it_a, it_b, h_a[it_b])
    GAPS.append(abs(it_a - it_b))
    return abs(it_a - it_b)


def oracle_stop_range(A, b, tol, orderings=200, seed=0):
    """(min, max) stop of the float64 oracle (the reference GPU loop) over
    random symmetric reorderings of one system: how far rounding alone moves
    the stop of that system."""
    import oracle

    rng = np.random.default_rng(seed)
    A = A.tocsr()
    its = []
    for _ in range(orderings):
        p = rng.permutation(A.shape[0])
        _, it, _, _ = oracle.pcg_gpu_loop(A[p][:, p].tocsr(), np.asarray(b)[p], tol=tol)
        its.append(it)
    return min(its), max(its)


SPREAD_CASES: list = []  # (gap, oracle stop range, hip stop) where the spread rule applied


def check_stop_gap(it_a, it_b, system=None, tol=None):
    """Record and bound the stop gap of two solves of one system.

    At most MAX_ITER_GAP apart; where the gap is larger and the system
    (A, b) is given, it_a must lie within one iteration of the stops the
    float64 oracle itself takes over 200 random reorderings of that system
    (on some C3 systems those span 5-6 iterations)."""
    gap = abs(it_a - it_b)
    if gap > MAX_ITER_GAP and system is not None:
        lo, hi = oracle_stop_range(system[0], system[1], tol)
        assert lo - 1 <= it_a <= hi + 1, (it_a, it_b, (lo, hi))
        SPREAD_CASES.append((gap, (lo, hi), it_a))
    else:
        assert gap <= MAX_ITER_GAP, (it_a, it_b)
    GAPS.append(gap)
