"""GPU parity of the GMRES leg (``cc_gmres_jacobi_f64`` through
``fast_solver_backends.solve_gpu_gmres``) against the reference's CPU leg
``solve_cpu_gmres`` (SciPy gmres, ``code/fast_solver_backends.py:135-191``).

Systems: a nonsymmetric convection-diffusion operator on the reference's
5-degree oblique cut (``E^T (K + H) E + beta E^T Q_x E``, Q_x the x-derivative
form of the same assembly) and an SPD sliver system with a short restart.
Bars: the inner-step count within ``GMRES_ITER_GAP`` of SciPy's (both sum the
same recurrences in different orders) where the restart is long enough for
steady convergence; for GMRES(25) on the convection operator, which
stagnates between restarts, within the spread that reordering the same
system moves SciPy's own count (``STAGNATION_SPREAD``, measured in
``tests/test_oracle.py::test_restarted_gmres_count_moves_with_rounding``);
the true residual below the reference's success bar, the solution within 1e-6
of a sparse direct solve.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cutcell.basis import make_open_uniform_basis
from cutcell.oblique2d import (HalfPlaneCut, assemble_oblique_cut, polynomial_preserving_extension, pullback,
                               sliver_system, sym)

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GMRES_ITER_GAP = 3
STAGNATION_SPREAD = 0.35  # SciPy's GMRES(25) count over 20 reorderings: 1736-2644


def _convection_system(beta=20.0):
    bx = make_open_uniform_basis(32, 4)
    by = make_open_uniform_basis(32, 4)
    s = assemble_oblique_cut(bx, by, HalfPlaneCut.from_angle(np.deg2rad(5.0), 0.85 / 32), 24, with_sbp=True)
    agg = polynomial_preserving_extension(s, 0.05, 4)
    A = (sym(pullback(s.K, agg.E) + pullback(s.H, agg.E)) + beta * pullback(s.Qx, agg.E)).tocsr()
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    return A, b


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    from cutcell import hip

    hip.load_library()
    return hip


@pytest.mark.parametrize("case,restart", [("convection", 80), ("convection", 25), ("sliver", 20)])
def test_gmres_matches_scipy(gpu, case, restart):
    from fast_solver_backends import solve_cpu_gmres, solve_gpu_gmres

    if case == "convection":
        A, b = _convection_system()
        assert abs(A - A.T).max() > 0.05 * abs(A).max()  # genuinely nonsymmetric
    else:
        s = sliver_system(16, 4, 1e-3)
        A, b = s.A, s.b
    x_c, rc = solve_cpu_gmres(A, b, tol=1e-8, restart=restart)
    x_g, rg = solve_gpu_gmres(A, b, tol=1e-8, restart=restart)
    print(f"{case} restart {restart}: GPU {rg.iterations} it, SciPy {rc.iterations} it, "
          f"rel {rg.relative_residual:.2e}", flush=True)
    assert rc.success and rg.success and rg.method == "hip-gmres" and rg.backend == "gpu"
    gap = GMRES_ITER_GAP if restart != 25 else STAGNATION_SPREAD * rc.iterations
    assert abs(rg.iterations - rc.iterations) <= gap
    assert rg.relative_residual <= 1e-8 * 1.0001
    xs = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x_g - xs) / np.linalg.norm(xs) < 1e-6


def test_gmres_edge_cases(gpu):
    from fast_solver_backends import solve_gpu_gmres

    n = 9
    A = sp.diags(np.arange(1.0, n + 1.0), format="csr")
    x, rep = solve_gpu_gmres(A, np.zeros(n))
    assert rep.success and rep.iterations == 0 and not x.any()
    # A diagonal system is solved by the first Krylov vector (breakdown exit).
    b = np.random.default_rng(1).standard_normal(n)
    x, rep = solve_gpu_gmres(A, b, tol=1e-12)
    assert rep.success and rep.iterations == 1
    np.testing.assert_allclose(x, b / np.arange(1.0, n + 1.0), rtol=1e-13)
    # One restart of length 2 is not enough for a 40-unknown convection chain.
    m = 40
    C = sp.diags([-1.3 * np.ones(m - 1), 2.5 * np.ones(m), -0.7 * np.ones(m - 1)], [-1, 0, 1], format="csr")
    _, rep = solve_gpu_gmres(C, np.ones(m), tol=1e-12, restart=2, maxiter=1)
    assert not rep.success and rep.message.startswith("not converged") and rep.iterations == 2
