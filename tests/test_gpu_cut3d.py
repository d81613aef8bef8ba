"""GPU parity of the 3D cut-cell Gram blocks (csrc/cut3d.hip) against the
host integration (cutcell/grid3d.py levelset_cut_data): the same Kuhn
sub-tetrahedra and collapsed Gauss rule, so the blocks agree to the rounding
of their sums, and the matrix-free operator built from either applies the
same."""

import numpy as np
import pytest

from cutcell.cut3d import PlanarCut, SphereObstacle
from cutcell.grid3d import Axis, Grid3D, levelset_cut_data
from cutcell.workloads import bump_levelset

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = {
    "planar-p3-N8": (lambda: Grid3D((Axis(8, 3), Axis(8, 3), Axis(8, 3))), lambda: PlanarCut.random(3)),
    "sphere-p3-N10": (lambda: Grid3D((Axis(10, 3), Axis(10, 3), Axis(10, 3))), lambda: SphereObstacle.random(1)),
    "sphere-p2-N12": (lambda: Grid3D((Axis(12, 2), Axis(12, 2), Axis(12, 2))), lambda: SphereObstacle.random(2)),
    "sphere-p1-N9": (lambda: Grid3D((Axis(9, 1), Axis(9, 1), Axis(9, 1))), lambda: SphereObstacle.random(5)),
    "bump-periodic-p3": (
        lambda: Grid3D((Axis(16, 3, 9.0, True), Axis(9, 3, 3.036, False), Axis(8, 3, 4.5, True))),
        lambda: bump_levelset(),
    ),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    gfn, lfn = CASES[request.param]
    grid, lev = gfn(), lfn()
    return request.param, grid, lev, levelset_cut_data(grid, lev, jobs=1)


def test_gram_blocks_match_host(case):
    """Same cut cells and classes; volume fractions to 1e-13; each block to
    1e-12 of its largest entry; gamma to 1e-12. The device blocks are exactly
    symmetric before the common symmetrisation."""
    name, grid, lev, host = case
    dev = levelset_cut_data(grid, lev, device=True)
    np.testing.assert_array_equal(dev.cell_class, host.cell_class)
    np.testing.assert_array_equal(dev.cut_cells, host.cut_cells)
    n = host.cut_cells.shape[0]
    assert n > 0
    assert np.abs(dev.cut_fraction - host.cut_fraction).max() <= 1e-13
    for blk in ("cut_mass", "cut_stiff"):
        a, b = getattr(dev, blk), getattr(host, blk)
        scale = np.abs(b).reshape(n, -1).max(axis=1)[:, None, None]
        err = (np.abs(a - b) / scale).max()
        print(f"{name} {blk}: {n} cut cells, max err {err:.2e} of the block's largest entry", flush=True)
        assert err <= 1e-12
    assert np.abs(dev.gamma - host.gamma).max() <= 1e-12


def test_device_blocks_are_exactly_symmetric(case):
    from cutcell.grid3d import gram_blocks_device

    name, grid, lev, host = case
    ax, ay, az = grid.axes
    nex, ney, nez = grid.ne
    xs, ys, zs = np.arange(nex + 1) * ax.h, np.arange(ney + 1) * ay.h, np.arange(nez + 1) * az.h
    Zc, Yc, Xc = np.meshgrid(zs, ys, xs, indexing="ij")
    phi = lev(Xc, Yc, Zc)
    cells = host.cut_cells
    corner = np.stack([phi[cells[:, 2] + ((c >> 2) & 1), cells[:, 1] + ((c >> 1) & 1), cells[:, 0] + (c & 1)]
                       for c in range(8)], axis=1)
    q = grid.degree + 3
    m, st, fr = gram_blocks_device(grid, cells, corner, q)
    assert np.array_equal(m, m.transpose(0, 2, 1))
    assert np.array_equal(st, st.transpose(0, 2, 1))


def test_matrix_free_operator_from_device_blocks(case):
    """The matrix-free system built with the device integration has the same
    extension E (pattern and values to 1e-12) and applies as the host-built
    one to 1e-11 of |A| |x|."""
    from cutcell.matfree import DeviceMatrixFree, assemble_reduced_csr, build_matrix_free

    name, grid, lev, host = case
    if grid.degree < 2:
        pytest.skip("the matrix-free operator is built for degrees 2 and 3")
    mass = 0.0 if "bump" in name else 1.0
    mh = build_matrix_free(grid, lev, mass_coef=mass)
    md = build_matrix_free(grid, lev, mass_coef=mass, device=True)
    np.testing.assert_array_equal(md.interior, mh.interior)
    np.testing.assert_array_equal(md.small, mh.small)
    Eh, Ed = mh.E.tocsr(), md.E.tocsr()
    Eh.sort_indices()
    Ed.sort_indices()
    np.testing.assert_array_equal(Ed.indptr, Eh.indptr)
    np.testing.assert_array_equal(Ed.indices, Eh.indices)
    assert np.abs(Ed.data - Eh.data).max() <= 1e-12 * max(np.abs(Eh.data).max(), 1.0)
    A = assemble_reduced_csr(mh)
    x = np.random.default_rng(9).standard_normal(mh.n)
    yh = DeviceMatrixFree(mh).apply(torch.from_numpy(x).to("cuda")).cpu().numpy()
    yd = DeviceMatrixFree(md).apply(torch.from_numpy(x).to("cuda")).cpu().numpy()
    assert np.abs(yd - yh).max() <= 1e-11 * (abs(A) @ np.abs(x)).max()
