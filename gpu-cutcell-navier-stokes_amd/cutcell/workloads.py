"""Seeded system recipes for the five benchmark configurations.

- ``c1_system``: 2D Poisson, p = 4, 32x32 elements, straight cut at 5 degrees
  with offset 0.85 h, manufactured ``u = sin(pi x) sin(pi y)`` (screened
  operator ``-lap u + u``), the reference's oblique-cut recipe
  (``code/oblique_cut_sbp.py:271-433, 614-672``).
- ``c2_system``: 3D cut box, p = 3, planar cut with seeded random normal and
  offset, grid sized to ~81k aggregated unknowns (``cutcell.cut3d``).
- ``c3_batch``: 72 oblique cuts (angle U[0, 45] degrees, offset U[0.05, 0.95] h)
  times p in {3, 4, 5} on 48x48 elements, plus 25 p = 4 slivers with cut-cell
  volume fraction 10^U[-5, -1] (the sliver recipe of
  ``code/fast_solver_workbench.py:138-168``).
- ``c4_assembly``: 3D box, p = 3, 64^3 elements, sphere obstacle with radius
  U[0.2, 0.4] and a centre keeping the ball inside.
"""

from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .oblique2d import oblique_system, sliver_system


def c1_system():
    return oblique_system(32, 4, np.deg2rad(5.0), 0.85 / 32, rhs="manufactured")


def c2_system(seed: int = 0, target: int = 81_000):
    from .cut3d import planar_cut_system

    return planar_cut_system(seed=seed, degree=3, target_unknowns=target)


def _c3_case(args):
    kind, n_el, p, a, b, seed = args
    if kind == "oblique":
        s = oblique_system(n_el, p, np.deg2rad(a), b / n_el, rhs="gaussian", seed=seed)
        s.meta.update(angle_deg=a, offset_fraction=b)
    else:
        s = sliver_system(n_el, p, a)
    return s


def c3_cases(n_elements: int = 48, n_cuts: int = 72, degrees=(3, 4, 5), n_slivers: int = 25, seed: int = 0):
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0.0, 45.0, n_cuts)
    offsets = rng.uniform(0.05, 0.95, n_cuts)
    eps = 10.0 ** rng.uniform(-5.0, -1.0, n_slivers)
    cases = []
    idx = 0
    for i in range(n_cuts):
        for p in degrees:
            cases.append(("oblique", n_elements, p, float(angles[i]), float(offsets[i]), 1000 * seed + idx))
            idx += 1
    for e in eps:
        cases.append(("sliver", n_elements, 4, float(e), 0.0, 0))
    return cases


def c3_batch(n_elements: int = 48, n_cuts: int = 72, degrees=(3, 4, 5), n_slivers: int = 25, seed: int = 0,
             jobs: int | None = None):
    """Build the batched configuration's systems (process pool over cases)."""
    cases = c3_cases(n_elements, n_cuts, degrees, n_slivers, seed)
    jobs = jobs or min(16, os.cpu_count() or 1)
    if jobs <= 1 or len(cases) < 4:
        return [_c3_case(c) for c in cases]
    with ProcessPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(_c3_case, cases, chunksize=4))
