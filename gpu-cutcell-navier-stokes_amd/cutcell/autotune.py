"""Project autotuner for the HIP launch shapes.

Two tables, cached on disk as JSON (``$CUTCELL_AUTOTUNE_CACHE`` or
``~/.cache/cutcell/autotune.json``) and keyed by GPU architecture and a bucket
of the problem shape:

- ``spmm``: for one matrix shape class and k right-hand sides, the nonzero
  lanes per row team, the grid and the unroll depth of the CSR SpMM that the
  CG loop runs every iteration (set in the library with
  ``cc_spmm_set_config``);
- ``batched``: for a batch of small systems, the workgroup size and the lanes
  per row team of the persistent batched solve.

Candidates are timed with HIP events on the current stream; the fastest is
kept. Timing runs only short, bounded work (a few SpMM launches, or a batched
solve capped at a few tens of iterations).
"""

from __future__ import annotations

import ctypes
import json
import math
import os
from pathlib import Path

from . import hip

_MEM: dict = {}

# unroll value that selects the nonzero-window SpMV kernel (k = 1).
WINDOW_UNROLL = 16


def cache_path() -> Path:
    p = os.environ.get("CUTCELL_AUTOTUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "cutcell" / "autotune.json"


def _load() -> dict:
    if _MEM:
        return _MEM
    p = cache_path()
    if p.exists():
        try:
            _MEM.update(json.loads(p.read_text()))
        except (OSError, ValueError):
            pass
    return _MEM


def _store(key: str, value: dict) -> None:
    table = _load()
    table[key] = value
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(table, indent=1, sort_keys=True))
    except OSError:
        pass


def _arch() -> str:
    import torch

    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(props, "gcnArchName", props.name).split(":")[0]


def _bucket(x: float) -> int:
    return int(round(math.log2(max(x, 1.0))))


def _time_ms(fn, reps: int) -> float:
    import torch

    s = torch.cuda.current_stream()
    fn()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(s)
    for _ in range(reps):
        fn()
    end.record(s)
    end.synchronize()
    return start.elapsed_time(end) / reps


# ---------------------------------------------------------------------------
# SpMM of the single-system CG loop.
# ---------------------------------------------------------------------------


def spmm_candidates(k: int, mean_row: float, value_bytes: int = 8, sweeps=(1, 8, 32)):
    """Launch shapes: vector width (right-hand sides per lane), nonzero slots
    per row team, grid, unroll depth, nontemporal matrix loads and sweep steps
    (chunks per block, taken grid-strided)."""
    cpls = [c for c in (1, 2, 4) if c <= k and c * value_bytes <= 16]
    for sweep in sweeps:
        for cpl in cpls:
            kl = k // cpl
            nzs = [nz for nz in (64, 32, 16, 8, 4, 2, 1) if kl * nz <= 64 and nz <= max(mean_row, 1)]
            nzs = nzs[:2] if nzs else [max(1, 64 // kl)]
            for nz in nzs:
                for grid in (512, 1024):
                    for unroll, stream in ((1, 0), (4, 0), (4, 1), (8, 0), (8, 1)):
                        yield dict(cpl=cpl, nz=nz, grid=grid, unroll=unroll, stream=stream, sweep=sweep)
        if k == 1:
            # The nonzero-window kernel (row shape does not matter to it).
            for grid in (512, 768, 1024):
                for stream in (0, 1):
                    yield dict(cpl=1, nz=64, grid=grid, unroll=WINDOW_UNROLL, stream=stream, sweep=sweep)


def spmm_config(A: "hip.DeviceCSR", k: int, precision: str = "fp64", force: bool = False) -> dict:
    """Pick and apply the SpMM launch shape for ``A`` with ``k`` right-hand sides."""
    import torch

    lib = hip.require_gpu()
    mean_row = A.nnz / max(A.n, 1)
    vb = 8 if precision == "fp64" else 4
    key = f"spmm5|{_arch()}|{precision}|k{k}|n{_bucket(A.n)}|row{_bucket(mean_row)}"
    table = _load()
    if key in table and not force:
        cfg = table[key]
        lib.cc_spmm_set_config(k, vb, cfg["nz"], cfg["grid"], cfg["unroll"], cfg["cpl"], cfg.get("stream", 0),
                               cfg.get("sweep", 0))
        return cfg
    dt = torch.float64 if precision == "fp64" else torch.float32
    X = torch.randn(A.n, k, dtype=dt, device=A.vals.device)
    Y = torch.empty_like(X)
    vals = A.vals if precision == "fp64" else A.ensure_f32()
    fn_name = "cc_spmm_csr_f64" if precision == "fp64" else "cc_spmm_csr_f32"
    fn = getattr(lib, fn_name)
    best = None
    for cand in spmm_candidates(k, mean_row, vb):
        if lib.cc_spmm_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"], cand["stream"],
                                  cand["sweep"]) != 0:
            continue
        call = lambda: fn(A.n, k, hip._ptr(A.indptr), hip._ptr(A.indices), hip._ptr(vals), hip._ptr(X), hip._ptr(Y),
                          hip._stream())
        ms = _time_ms(call, 5)
        if best is None or ms < best[0]:
            best = (ms, cand)
    cfg = dict(best[1], ms=best[0])
    lib.cc_spmm_set_config(k, vb, cfg["nz"], cfg["grid"], cfg["unroll"], cfg["cpl"], cfg["stream"], cfg["sweep"])
    _store(key, cfg)
    return cfg


def rb_candidates(k: int, value_bytes: int = 8):
    """Row-block product shapes: vector width, union-column slots per group
    team, grid and unroll depth."""
    for cpl in [c for c in (1, 2, 4) if c <= k and c * value_bytes <= 16]:
        kl = k // cpl
        for nz in [nz for nz in (16, 8, 4, 2) if kl * nz <= 64]:
            for grid in (512, 1024):
                for unroll in (1, 4):
                    yield dict(cpl=cpl, nz=nz, grid=grid, unroll=unroll)


def rb_config(rb: "hip.DeviceRowBlocks", k: int, precision: str = "fp64", force: bool = False) -> dict:
    """Pick and apply the row-block product shape for ``rb`` with ``k``
    right-hand sides."""
    import torch

    lib = hip.require_gpu()
    vb = 8 if precision == "fp64" else 4
    mean_cols = rb.nu / max(rb.ngroups, 1)
    key = f"rb1|{_arch()}|{precision}|k{k}|n{_bucket(rb.n)}|cols{_bucket(mean_cols)}"
    table = _load()
    if key in table and not force:
        cfg = table[key]
        lib.cc_rb_set_config(k, vb, cfg["nz"], cfg["grid"], cfg["unroll"], cfg["cpl"])
        return cfg
    dt = torch.float64 if precision == "fp64" else torch.float32
    X = torch.randn(rb.n, k, dtype=dt, device=rb.gptr.device)
    Y = torch.empty_like(X)
    fn = lib.cc_spmm_rb_f64 if precision == "fp64" else lib.cc_spmm_rb_f32
    best = None
    for cand in rb_candidates(k, vb):
        if lib.cc_rb_set_config(k, vb, cand["nz"], cand["grid"], cand["unroll"], cand["cpl"]) != 0:
            continue
        call = lambda: fn(ctypes.addressof(rb.desc), k, hip._ptr(X), hip._ptr(Y), hip._stream())
        ms = _time_ms(call, 5)
        if best is None or ms < best[0]:
            best = (ms, cand)
    cfg = dict(best[1], ms=best[0])
    lib.cc_rb_set_config(k, vb, cfg["nz"], cfg["grid"], cfg["unroll"], cfg["cpl"])
    _store(key, cfg)
    return cfg


def reset_spmm_config() -> None:
    lib = hip.load_library()
    for k in hip.SUPPORTED_K:
        for vb in (8, 4):
            lib.cc_spmm_set_config(k, vb, 0, 0, 0, 0, 0, 0)
            lib.cc_rb_set_config(k, vb, 0, 0, 0, 0)


# ---------------------------------------------------------------------------
# Batched persistent solve.
# ---------------------------------------------------------------------------


def batched_candidates(n_max: int):
    for block in (256, 512, 1024):
        rpt = -(-n_max // block)
        if rpt > 16 or (rpt > 8 and block * 16 > 8192):
            continue
        for team in (8, 16, 32, 64):
            yield dict(block_threads=block, team_lanes=team)


def batched_config(bs, precision: str = "fp64", force: bool = False, tune_iters: int = 30) -> dict:
    """Pick the workgroup size and row-team width for a batch of systems."""
    import torch

    key = f"batched|{_arch()}|{precision}|nmax{_bucket(bs.n_max)}|row{_bucket(bs.mean_row)}|sys{_bucket(bs.nsys)}"
    table = _load()
    if key in table and not force:
        return table[key]
    B = torch.ones(bs.total_n, dtype=torch.float64, device=bs.device)
    X = torch.empty_like(B)
    best = None
    for cand in batched_candidates(bs.n_max):
        call = lambda: bs.solve_device(B, X, tol=0.0, maxiter=tune_iters, precision=precision, **cand)
        try:
            ms = _time_ms(call, 2)
        except hip.HipError:
            continue
        if best is None or ms < best[0]:
            best = (ms, cand)
    if best is None:
        raise hip.HipError("no batched launch shape could be run")
    cfg = dict(best[1], ms_per_iter=best[0] / tune_iters)
    _store(key, cfg)
    return cfg
