"""A versioned on-disk format for aggregated cut-cell systems.

One ``.npz`` file (NumPy, no pickled objects) holds a system and its
aggregation, so systems built here and systems exported from the reference
(``code/bell_3d_core.py``, ``code/oblique_cut_sbp.py``) can be stored side by
side and compared with :func:`diff_systems`.

Keys (version 1, ``format = "cutcell-aggregated-system"``):

- ``format``, ``version``;
- ``meta_json``: a JSON object (kind, degree, N, cut parameters, ...);
- ``A_shape``, ``A_indptr``, ``A_indices``, ``A_data``: the SPD operator
  ``sym(E^T K E + E^T H E)`` in CSR (``A_data`` may be absent: pattern only);
- ``E_shape``, ``E_indptr``, ``E_indices``, ``E_data``: the extension ``E``
  (``n_full x n_reduced``) in CSR;
- ``active``, ``interior``, ``small``: integer index maps into the full
  background dofs; ``gamma``: the cut-fraction indicator per active dof;
- ``b``: a right-hand side (optional).

Version 0 is the layout of the committed aggregation fixtures under
``tests/golden`` (written before this module existed): the same array keys,
no ``format`` / ``version`` / ``meta_json``; the 3D fixtures store the
operator's pattern only (no ``A_data``). :func:`load_system` reads both.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp

FORMAT = "cutcell-aggregated-system"
VERSION = 1


class FormatError(ValueError):
    """The file is not an aggregated system of a version this module reads."""


@dataclass
class AggregatedSystem:
    A: sp.csr_matrix | None  # None when only the pattern is stored
    A_pattern: sp.csr_matrix | None  # ones on A's pattern
    E: sp.csr_matrix | None
    active: np.ndarray | None
    interior: np.ndarray | None
    small: np.ndarray | None
    gamma: np.ndarray | None
    b: np.ndarray | None
    meta: dict = field(default_factory=dict)
    version: int = VERSION


def _csr_arrays(prefix: str, M) -> dict:
    M = sp.csr_matrix(M)
    M.sort_indices()
    return {f"{prefix}_shape": np.asarray(M.shape, dtype=np.int64), f"{prefix}_indptr": M.indptr.astype(np.int64),
            f"{prefix}_indices": M.indices.astype(np.int32), f"{prefix}_data": M.data.astype(np.float64)}


def save_system(path, *, A=None, E=None, b=None, active=None, interior=None, small=None, gamma=None,
                meta: dict | None = None) -> Path:
    """Write a version-1 file; every part is optional."""
    out = {"format": np.asarray(FORMAT), "version": np.asarray(VERSION, dtype=np.int64),
           "meta_json": np.asarray(json.dumps(meta or {}, sort_keys=True))}
    if A is not None:
        out.update(_csr_arrays("A", A))
    if E is not None:
        out.update(_csr_arrays("E", E))
    for key, arr, dt in (("active", active, np.int64), ("interior", interior, np.int64), ("small", small, np.int64),
                         ("gamma", gamma, np.float64), ("b", b, np.float64)):
        if arr is not None:
            out[key] = np.asarray(arr, dtype=dt)
    path = Path(path)
    with path.open("wb") as f:
        np.savez_compressed(f, **out)
    return path


def _csr(d, prefix: str, with_data: bool):
    if f"{prefix}_indptr" not in d:
        return None
    shape = tuple(int(v) for v in d[f"{prefix}_shape"])
    indptr = d[f"{prefix}_indptr"]
    indices = d[f"{prefix}_indices"]
    data = d[f"{prefix}_data"] if with_data else np.ones(indices.size)
    return sp.csr_matrix((data, indices, indptr), shape=shape)


def load_system(path) -> AggregatedSystem:
    """Read a version-1 file or a version-0 aggregation fixture."""
    with np.load(Path(path), allow_pickle=False) as d:
        keys = set(d.files)
        if "format" in keys:
            if str(d["format"]) != FORMAT:
                raise FormatError(f"{path}: format {str(d['format'])!r}, expected {FORMAT!r}")
            version = int(d["version"])
            if version != VERSION:
                raise FormatError(f"{path}: version {version} is not readable (this module reads 0 and {VERSION})")
            meta = json.loads(str(d["meta_json"]))
        elif "E_indptr" in keys and "A_indptr" in keys:
            version, meta = 0, {}
        else:
            raise FormatError(f"{path}: not an aggregated system file (no format key and no fixture arrays)")
        has_a_data = "A_data" in keys
        A = _csr(d, "A", True) if has_a_data else None
        A_pattern = _csr(d, "A", False)
        E = _csr(d, "E", True)
        get = lambda k: d[k].copy() if k in keys else None
        return AggregatedSystem(A=A, A_pattern=A_pattern, E=E, active=get("active"), interior=get("interior"),
                                small=get("small"), gamma=get("gamma"), b=get("b"), meta=meta, version=version)


def _pattern_equal(X, Y) -> bool:
    if X is None or Y is None:
        return X is None and Y is None
    X = sp.csr_matrix(X)
    Y = sp.csr_matrix(Y)
    X.sort_indices()
    Y.sort_indices()
    return X.shape == Y.shape and np.array_equal(X.indptr, Y.indptr) and np.array_equal(X.indices, Y.indices)


def diff_systems(a: AggregatedSystem, b: AggregatedSystem) -> dict:
    """Integer parts compared exactly (index maps, sparsity of E and A);
    values by their largest relative difference on a shared pattern."""
    out = {}
    for key in ("active", "interior", "small"):
        x, y = getattr(a, key), getattr(b, key)
        out[f"{key}_equal"] = None if x is None or y is None else bool(np.array_equal(x, y))
    out["E_pattern_equal"] = _pattern_equal(a.E, b.E)
    out["A_pattern_equal"] = _pattern_equal(a.A_pattern, b.A_pattern)
    if out["E_pattern_equal"] and a.E is not None:
        out["E_max_rel_diff"] = float(np.abs(a.E.data - b.E.data).max(initial=0.0) / max(np.abs(a.E.data).max(initial=0.0), 1e-300))
    if out["A_pattern_equal"] and a.A is not None and b.A is not None:
        A1, A2 = sp.csr_matrix(a.A), sp.csr_matrix(b.A)
        A1.sort_indices()
        A2.sort_indices()
        out["A_max_rel_diff"] = float(np.abs(A1.data - A2.data).max(initial=0.0) / max(np.abs(A1.data).max(initial=0.0), 1e-300))
    return out


def from_cut_system(s) -> AggregatedSystem:
    """The in-memory form of a ``CutSystem`` / ``System3D`` built here."""
    A = sp.csr_matrix(s.A)
    A.sort_indices()
    P = A.copy()
    P.data = np.ones_like(P.data)
    return AggregatedSystem(A=A, A_pattern=P, E=sp.csr_matrix(s.E), active=np.asarray(s.active),
                            interior=np.asarray(s.interior), small=np.asarray(s.small_support), gamma=None,
                            b=np.asarray(getattr(s, "b", None)) if getattr(s, "b", None) is not None else None,
                            meta=dict(getattr(s, "meta", {})))


def save_cut_system(path, s) -> Path:
    return save_system(path, A=s.A, E=s.E, b=getattr(s, "b", None), active=s.active, interior=s.interior,
                       small=s.small_support, meta={k: v for k, v in getattr(s, "meta", {}).items()
                                                     if isinstance(v, (int, float, str, bool, list))})
