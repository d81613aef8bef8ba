"""A pressure-projection time loop whose solves run on the GPU.

The reference's flow drivers factor their coupled velocity-pressure systems
once with SuperLU and back-substitute every step
(``code/channel_cylinder_flow.py:1128-1139``,
``code/tg3d_fully_conservative.py:570-599``). This module is the same kind of
consumer built on the aggregated cut-cell operators of this package, with
every solve an SPD Jacobi-PCG through the HIP library:

    u* = u^n + dt F(t_n)                    forcing in the velocity space
    L phi = b,  b_j = sum_d (G_d^T u*_d)_j   pressure Poisson (L = E^T K E,
                                             natural boundary; b sums to 0)
    M u^{n+1}_d = M u*_d - G_d phi           velocity correction (M = E^T H E,
                                             both components in one solve)

with G_d = E^T Q_d E, Q_d[i, j] = int phi_i d_d phi_j on the cut domain (the
forms of ``code/oblique_cut_sbp.py:271-375``). ``L`` is the Laplacian, not
``G^T M^-1 G``: an approximate projection, which removes most but not all of
the discrete divergence (the tests check the reduction). ``GpuProjection``
runs the loop on the device (PCG, products and vector updates); the host
``HostProjection`` is its twin with SuperLU for ``M`` and SciPy CG for the
singular ``L``, to compare trajectories against.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .basis import make_open_uniform_basis
from .oblique2d import HalfPlaneCut, assemble_oblique_cut, polynomial_preserving_extension, pullback, sym


@dataclass
class ProjectionOperators:
    L: sp.csr_matrix  # pressure Poisson, E^T K E (singular: constants)
    M: sp.csr_matrix  # velocity mass, E^T H E (SPD)
    G: tuple  # (G_x, G_y), E^T Q_d E
    greville: np.ndarray  # (n, 2) coordinates of the reduced unknowns
    meta: dict = field(default_factory=dict)


def build_operators(n_elements: int = 32, degree: int = 3, angle_deg: float = 20.0, offset: float = 0.3):
    """The cut unit square n . x >= offset with its projection operators."""
    bx = make_open_uniform_basis(n_elements, degree)
    by = make_open_uniform_basis(n_elements, degree)
    cut = HalfPlaneCut.from_angle(np.deg2rad(angle_deg), offset)
    s = assemble_oblique_cut(bx, by, cut, max(24, 2 * degree + 14), with_sbp=True)
    agg = polynomial_preserving_extension(s, 0.05, degree)
    L = sym(pullback(s.K, agg.E))
    M = sym(pullback(s.H, agg.E))
    G = (pullback(s.Qx, agg.E).tocsr(), pullback(s.Qy, agg.E).tocsr())
    gx, gy = np.meshgrid(bx.greville, by.greville, indexing="xy")
    gre = np.stack([gx.ravel(), gy.ravel()], axis=1)[agg.interior]
    return ProjectionOperators(L, M, G, gre, dict(n_elements=n_elements, degree=degree, angle_deg=angle_deg,
                                                  offset=offset, n=L.shape[0]))


def forcing(ops: ProjectionOperators, t: float) -> np.ndarray:
    """A smooth forcing with gradient and rotational parts, as spline
    coefficients at the Greville points: (n, 2)."""
    x, y = ops.greville[:, 0], ops.greville[:, 1]
    fx = np.cos(np.pi * x) * np.sin(2 * np.pi * y) * (1.0 + 0.5 * np.sin(t)) + x * y
    fy = -np.sin(2 * np.pi * x) * np.cos(np.pi * y) * (1.0 + 0.5 * np.cos(t)) + x - y
    return np.stack([fx, fy], axis=1)


def divergence(ops: ProjectionOperators, u: np.ndarray) -> np.ndarray:
    """b_j = sum_d (G_d^T u_d)_j, the weak divergence (int u . grad phi_j)."""
    return ops.G[0].T @ u[:, 0] + ops.G[1].T @ u[:, 1]


@dataclass
class StepReport:
    t: float
    div_before: float
    div_after: float
    pressure_iterations: int
    velocity_iterations: int


class HostProjection:
    """The loop on the host: SuperLU for M (factored once), SciPy CG for L."""

    def __init__(self, ops: ProjectionOperators, dt: float = 1e-2, tol: float = 1e-10):
        from fast_solver_backends import solve_cpu_cg

        self.ops, self.dt, self.tol = ops, dt, tol
        self.lu = spla.splu(ops.M.tocsc())
        self._cg = solve_cpu_cg
        self.u = np.zeros((ops.L.shape[0], 2))
        self.t = 0.0

    def step(self) -> StepReport:
        ops = self.ops
        us = self.u + self.dt * forcing(ops, self.t)
        b = divergence(ops, us)
        phi, rep = self._cg(ops.L, b, tol=self.tol, maxiter=20000)
        rhs = np.stack([ops.M @ us[:, d] - ops.G[d] @ phi for d in range(2)], axis=1)
        self.u = np.stack([self.lu.solve(rhs[:, d]) for d in range(2)], axis=1)
        self.t += self.dt
        return StepReport(self.t, float(np.linalg.norm(b)), float(np.linalg.norm(divergence(ops, self.u))),
                          rep.iterations, 0)


class GpuProjection:
    """The loop on the device: every product, vector update and solve in HIP
    kernels (Jacobi-PCG for L on one column, for M on both components)."""

    def __init__(self, ops: ProjectionOperators, dt: float = 1e-2, tol: float = 1e-10):
        import torch

        from . import hip

        self.hip, self.torch = hip, torch
        self.ops, self.dt, self.tol = ops, dt, tol
        self.L = hip.DeviceCSR.from_scipy(ops.L)
        self.M = hip.DeviceCSR.from_scipy(ops.M)
        self.G = [hip.DeviceCSR.from_scipy(sp.csr_matrix(g)) for g in ops.G]
        self.GT = [hip.DeviceCSR.from_scipy(sp.csr_matrix(g.T)) for g in ops.G]
        self.u = torch.zeros((ops.L.shape[0], 2), dtype=torch.float64, device="cuda")
        self.t = 0.0

    def _div(self, u):
        hip = self.hip
        return hip.spmv(self.GT[0], u[:, 0].contiguous()) + hip.spmv(self.GT[1], u[:, 1].contiguous())

    def step(self) -> StepReport:
        hip, torch = self.hip, self.torch
        f = torch.from_numpy(forcing(self.ops, self.t)).to("cuda")
        us = self.u + self.dt * f
        b = self._div(us)
        phi, res = hip.pcg(self.L, b, tol=self.tol, maxiter=20000)
        if res[0].status != hip.CC_CONVERGED:
            raise RuntimeError(f"pressure solve: {hip.STATUS_MESSAGE[res[0].status]}")
        rhs = torch.stack([hip.spmv(self.M, us[:, d].contiguous()) - hip.spmv(self.G[d], phi) for d in range(2)],
                          dim=1).contiguous()
        u, resv = hip.pcg(self.M, rhs, tol=self.tol, maxiter=20000)
        if any(r.status != hip.CC_CONVERGED for r in resv):
            raise RuntimeError("velocity solve did not converge")
        self.u = u
        self.t += self.dt
        return StepReport(self.t, float(torch.linalg.norm(b)), float(torch.linalg.norm(self._div(u))),
                          int(res[0].iterations), int(max(r.iterations for r in resv)))
