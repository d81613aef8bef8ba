"""Test configuration: import paths and the ``gpu`` marker.

Like the reference's ``code/tests/conftest.py`` this puts the solver package
directory on ``sys.path`` so tests import ``fast_solver_backends`` and
``cutcell`` directly; the repository root is added for the ``oracle`` package.
"""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "gpu-cutcell-navier-stokes_amd"
for p in (str(PKG), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X, gfx950)")


def pytest_terminal_summary(terminalreporter):
    """Distribution of the stop gaps that parity.check_trajectory accepted."""
    try:
        import parity
    except ImportError:
        return
    if parity.GAPS:
        hist = [parity.GAPS.count(g) for g in range(max(parity.GAPS) + 1)]
        terminalreporter.write_line(f"parity stop gaps (count of |gap| = 0, 1, ...): {hist}, "
                                    f"largest {max(parity.GAPS)} of {len(parity.GAPS)} trajectories")
    for gap, rng, it in parity.SPREAD_CASES:
        terminalreporter.write_line(f"stop gap {gap} accepted: HIP stop {it}, oracle stops {rng[0]}..{rng[1]} "
                                    "over reorderings of the same system")
