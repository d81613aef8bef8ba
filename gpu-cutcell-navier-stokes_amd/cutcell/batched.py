"""Batched persistent solves of many uneven cut-cell systems in one launch.

Host side of ``cc_pcg_batched``: concatenates the systems' CSR arrays (local
indices), orders them longest-first (nnz times a sqrt(n) iteration proxy) so
the device work counter hands the most expensive systems out first, and
reports one ``PcgResult`` per system. The launch shape (workgroup size, lanes
per row team) comes from ``cutcell.autotune`` unless given.

This replaces a host loop of ``solve_spd`` calls over independent cases (the
reference runs its sweep cases one by one, ``code/fast_solver_workbench.py:
201-217``); every system follows the reference GPU loop's semantics
(``code/fast_solver_backends.py:233-274``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import hip

N_MAX = 8192


@dataclass
class BatchedResult:
    X: list
    results: list
    config: dict


class BatchedSystems:
    """Device-resident concatenation of independent CSR systems."""

    def __init__(self, mats, device: str = "cuda", with_f32: bool = True):
        import torch

        mats = [m.tocsr() for m in mats]
        if not mats:
            raise ValueError("at least one system is required")
        for m in mats:
            if m.shape[0] != m.shape[1]:
                raise ValueError("square systems are required")
            if m.shape[0] > N_MAX:
                raise ValueError(f"batched systems are limited to {N_MAX} rows (got {m.shape[0]})")
        self.nsys = len(mats)
        self.n = np.array([m.shape[0] for m in mats], dtype=np.int32)
        self.nnz = np.array([m.nnz for m in mats], dtype=np.int64)
        self.n_max = int(self.n.max())
        ip_off = np.zeros(self.nsys, dtype=np.int64)
        nnz_off = np.zeros(self.nsys, dtype=np.int64)
        vec_off = np.zeros(self.nsys, dtype=np.int64)
        ip_off[1:] = np.cumsum(self.n[:-1].astype(np.int64) + 1)
        nnz_off[1:] = np.cumsum(self.nnz[:-1])
        vec_off[1:] = np.cumsum(self.n[:-1].astype(np.int64))
        self.vec_off = vec_off
        cost = self.nnz * np.sqrt(np.maximum(self.n, 1))
        order = np.argsort(-cost, kind="stable").astype(np.int32)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
        self.d_n = t(self.n, np.int32)
        self.d_ip_off = t(ip_off, np.int64)
        self.d_nnz_off = t(nnz_off, np.int64)
        self.d_vec_off = t(vec_off, np.int64)
        self.d_order = t(order, np.int32)
        self.d_indptr = t(np.concatenate([m.indptr for m in mats]), np.int32)
        self.d_indices = t(np.concatenate([m.indices for m in mats]), np.int32)
        self.d_vals = t(np.concatenate([m.data for m in mats]), np.float64)
        self.d_vals32 = self.d_vals.to(torch.float32) if with_f32 else None
        self.d_counter = torch.zeros(1, dtype=torch.int32, device=device)
        self.d_results = torch.zeros(self.nsys * ctypes.sizeof(hip.PcgResult), dtype=torch.uint8, device=device)
        self.total_n = int(self.n.sum())
        self.total_nnz = int(self.nnz.sum())
        self.mean_row = float(self.total_nnz / max(self.total_n, 1))
        self.device = device

    def concat(self, vectors):
        import torch

        return torch.from_numpy(np.ascontiguousarray(np.concatenate(vectors), dtype=np.float64)).to(self.device)

    def split(self, X):
        Xh = X.cpu().numpy()
        return [Xh[o : o + n] for o, n in zip(self.vec_off, self.n)]

    def bytes_per_iteration(self, value_bytes: int = 8) -> int:
        """Algorithmic matrix bytes streamed by one CG iteration over all systems."""
        return int(self.total_nnz * (value_bytes + 4) + 4 * (self.total_n + self.nsys))

    def solve_device(self, B, X, *, tol=1.0e-8, maxiter=5000, preconditioner="jacobi", precision="fp64",
                     inner_tol=1.0e-4, max_outer=12, block_threads=None, team_lanes=None, grid=0):
        """Launch the persistent solve on the current stream (asynchronous)."""
        lib = hip.require_gpu()
        if block_threads is None or team_lanes is None:
            from .autotune import batched_config

            cfg = batched_config(self, precision)
            block_threads = block_threads or cfg["block_threads"]
            team_lanes = team_lanes or cfg["team_lanes"]
        prec = {"none": 0, "jacobi": 1}[preconditioner]
        mode = {"fp64": 0, "mixed": 1}[precision]
        if mode == 1 and self.d_vals32 is None:
            raise ValueError("mixed precision needs the float32 values (with_f32=True)")
        p = hip._ptr
        st = lib.cc_pcg_batched(
            self.nsys, self.n_max, p(self.d_n), p(self.d_ip_off), p(self.d_nnz_off), p(self.d_vec_off),
            p(self.d_order), p(self.d_indptr), p(self.d_indices), p(self.d_vals),
            p(self.d_vals32) if self.d_vals32 is not None else None, p(B), p(X), float(tol), int(maxiter), prec,
            mode, float(inner_tol), int(max_outer), int(block_threads), int(team_lanes), int(grid),
            p(self.d_counter), p(self.d_results), hip._stream(),
        )
        hip._check(st, "pcg_batched")
        return dict(block_threads=int(block_threads), team_lanes=int(team_lanes), precision=precision)

    def read_results(self):
        raw = self.d_results.cpu().numpy().tobytes()
        arr = (hip.PcgResult * self.nsys).from_buffer_copy(raw)
        return list(arr)

    def solve(self, rhs, **kw) -> BatchedResult:
        """Solve every system; ``rhs`` is a list of per-system vectors."""
        import torch

        B = self.concat(rhs)
        X = torch.empty_like(B)
        cfg = self.solve_device(B, X, **kw)
        torch.cuda.synchronize()
        return BatchedResult(X=self.split(X), results=self.read_results(), config=cfg)
