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
- ``c4_operator``: 3D box, p = 3, 64^3 elements, sphere obstacle with radius
  U[0.2, 0.4] and a centre keeping the ball inside (matrix-free).
- ``c5_operator``: channel 128 x 64 x 64 elements, p = 3, periodic in x and z
  (extents 9 x 3.036 x 4.5), lower wall cut by a cosine bump of height 1 over
  length 4 (matrix-free; ``mass_coef = 0`` for pressure projection).
"""

from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .oblique2d import oblique_system, sliver_system


def bump_levelset(x0: float = 2.5, length: float = 4.0, height: float = 1.0):
    """Fluid above a smooth cosine bump ``y >= h/2 (1 - cos(2 pi (x - x0)/L))``
    on ``x0 <= x <= x0 + L`` (flat wall ``y = 0`` elsewhere)."""

    def phi(x, y, z):
        inside = (x >= x0) & (x <= x0 + length)
        wall = np.where(inside, 0.5 * height * (1.0 - np.cos(2.0 * np.pi * (x - x0) / length)), 0.0)
        return y - wall

    return phi


def c4_grid(n_elements: int = 64, degree: int = 3):
    from .grid3d import Axis, Grid3D

    return Grid3D(tuple(Axis(n_elements, degree) for _ in range(3)))


def c4_operator(seed: int = 0, n_elements: int = 64, degree: int = 3, device: bool = False):
    """Sphere-obstacle box, matrix-free (``cutcell.matfree``); ``device``
    integrates the cut cells on the GPU."""
    from .cut3d import SphereObstacle
    from .matfree import build_matrix_free

    lev = SphereObstacle.random(seed)
    mf = build_matrix_free(c4_grid(n_elements, degree), lev, device=device)
    mf.meta.update(kind="sphere3d", seed=seed, center=lev.center.tolist(), radius=lev.radius)
    return mf


def c5_grid(nx: int = 128, ny: int = 64, nz: int = 64, degree: int = 3, lengths=(9.0, 3.036, 4.5)):
    """Channel box periodic in x and z, clamped in y (periodic-hill extents)."""
    from .grid3d import Axis, Grid3D

    return Grid3D((Axis(nx, degree, lengths[0], True), Axis(ny, degree, lengths[1], False),
                   Axis(nz, degree, lengths[2], True)))


def c5_operator(nx: int = 128, ny: int = 64, nz: int = 64, degree: int = 3, mass_coef: float = 1.0,
                device: bool = False):
    """Bump-channel operator ``E^T (K + c M) E``, matrix-free; ``mass_coef = 0``
    is the pressure-projection operator ``E^T K E`` (singular: the constants
    are its null space, so right-hand sides must sum to zero)."""
    from .matfree import build_matrix_free

    mf = build_matrix_free(c5_grid(nx, ny, nz, degree), bump_levelset(), mass_coef=mass_coef, device=device)
    mf.meta.update(kind="bump_channel", shape=[nx, ny, nz], mass_coef=mass_coef)
    return mf


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
