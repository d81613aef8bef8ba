"""The host twin of the pressure-projection loop (``cutcell.projection``):
its operators have the structure the scheme needs (L 1 = 0, G_d 1 = 0, so
every divergence right-hand side is compatible with the singular L), and each
step removes most of the discrete divergence the forcing adds."""

import numpy as np

from cutcell.projection import HostProjection, build_operators, divergence


def test_operators_and_host_loop():
    ops = build_operators(n_elements=16, degree=3)
    one = np.ones(ops.L.shape[0])
    assert np.abs(ops.L @ one).max() < 1e-13 * abs(ops.L).max()
    for G in ops.G:
        assert np.abs(G @ one).max() < 1e-14 * abs(G).max()
    u = np.random.default_rng(0).standard_normal((ops.L.shape[0], 2))
    assert abs(divergence(ops, u).sum()) < 1e-12 * np.abs(divergence(ops, u)).sum()
    loop = HostProjection(ops, dt=0.01)
    for _ in range(4):
        r = loop.step()
        assert r.div_after < 2e-3 * r.div_before
        assert r.pressure_iterations > 0
