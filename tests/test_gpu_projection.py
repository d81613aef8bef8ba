"""The pressure-projection loop on the GPU (``cutcell.projection.GpuProjection``:
every product, update and solve through the HIP library) against its host
twin (SuperLU for the mass matrix, SciPy CG for the singular Laplacian):
after 10 steps the velocity fields agree to 1e-7 relative (both solve to
relative residual 1e-10), and every step removes the divergence as the host's."""

import numpy as np
import pytest

from cutcell.projection import GpuProjection, HostProjection, build_operators

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_gpu_loop_follows_the_host_loop():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible ROCm GPU")
    ops = build_operators(n_elements=32, degree=3)
    host, dev = HostProjection(ops, dt=0.01), GpuProjection(ops, dt=0.01)
    for step in range(10):
        rh, rd = host.step(), dev.step()
        print(f"step {step}: div {rd.div_before:.3e} -> {rd.div_after:.3e} (host {rh.div_after:.3e}), "
              f"PCG {rd.pressure_iterations} / {rd.velocity_iterations} it", flush=True)
        assert rd.div_after < 2e-3 * rd.div_before
        assert rd.div_after == pytest.approx(rh.div_after, rel=1e-3, abs=1e-12)
    u_dev = dev.u.cpu().numpy()
    assert np.linalg.norm(u_dev - host.u) / np.linalg.norm(host.u) < 1e-7
