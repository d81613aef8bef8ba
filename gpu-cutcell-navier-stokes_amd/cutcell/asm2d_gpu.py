"""2D half-plane cut-cell assembly on the GPU (``csrc/asm2d.hip``).

``assemble_oblique_cut_gpu`` returns the same ``ObliqueCutSystem2D`` (mass
``H`` and stiffness ``K``) as the host ``assemble_oblique_cut`` without
boundary forms or loads, with the per-cell work on the device: the Duffy
Gram blocks of the cut cells, then the CSR matrices gathered from the cells
that cover each entry. The host keeps the integer classification of the
cells (the same corner test as the host assembler), the 1D element forms
and the background mass diagonal. ``assemble_batch_gpu`` does many systems
of one (N, p, q) with a single Gram launch for all their cut cells.
``oblique_batch_gpu``, ``sliver_system_gpu`` and ``c3_batch_gpu`` are the
host recipes of ``cutcell.oblique2d`` / ``cutcell.workloads`` with that
assembly.

Reference behaviour: ``code/oblique_cut_sbp.py:271-433`` (assembly),
``:575-672`` (active sets and E, run on the host as in ``cutcell.oblique2d``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import hip
from .basis import gauss_legendre, make_open_uniform_basis
from .oblique2d import (CutSystem, HalfPlaneCut, ObliqueCutSystem2D, _element_forms, polynomial_preserving_extension,
                        pullback, sym)


class Cut2DCell(ctypes.Structure):
    """``cc_cut2d_cell`` of the C header."""

    _fields_ = [("ix", ctypes.c_int32), ("iy", ctypes.c_int32), ("normal_x", ctypes.c_double),
                ("normal_y", ctypes.c_double), ("offset", ctypes.c_double)]


def _classify(N: int, cut: HalfPlaneCut):
    """Full and cut cells by the host assembler's corner test."""
    h = 1.0 / N
    xs = np.arange(N + 1) * h
    corner_in = (cut.normal[0] * xs[None, :] + cut.normal[1] * xs[:, None]) >= cut.offset - 1.0e-14
    all_in = corner_in[:-1, :-1] & corner_in[:-1, 1:] & corner_in[1:, :-1] & corner_in[1:, 1:]
    any_in = corner_in[:-1, :-1] | corner_in[:-1, 1:] | corner_in[1:, :-1] | corner_in[1:, 1:]
    return all_in, any_in & ~all_in


def assemble_batch_gpu(basis, cuts, quadrature_order: int) -> list[ObliqueCutSystem2D]:
    """``H`` and ``K`` of one system per cut, all on ``basis`` x ``basis``."""
    import torch

    lib = hip.require_gpu()
    p, N = basis.degree, basis.n_elements
    p1, nb = p + 1, basis.n_basis
    q = int(quadrature_order)
    dev = "cuda"
    classes = [_classify(N, c) for c in cuts]
    cells = []
    for c, (_, cutm) in zip(cuts, classes):
        cy, cx = np.nonzero(cutm)
        for iy, ix in zip(cy.tolist(), cx.tolist()):
            cells.append((ix, iy, float(c.normal[0]), float(c.normal[1]), float(c.offset)))
    ncut = len(cells)
    nloc2 = (p1 * p1) ** 2
    g, w = gauss_legendre(q)
    gx = torch.from_numpy(np.ascontiguousarray(g)).to(dev)
    gw = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    Hb = torch.zeros(max(ncut, 1) * nloc2, dtype=torch.float64, device=dev)
    Kb = torch.zeros_like(Hb)
    areas = torch.zeros(max(ncut, 1), dtype=torch.float64, device=dev)
    if ncut:
        arr = (Cut2DCell * ncut)(*[Cut2DCell(*c) for c in cells])
        raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
        hip._check(lib.cc_cut2d_gram(ncut, hip._ptr(raw), N, p, q, hip._ptr(gx), hip._ptr(gw), hip._ptr(Hb),
                                     hip._ptr(Kb), hip._ptr(areas), hip._stream()), "cut2d_gram")
    area_h = areas.cpu().numpy()[:ncut]
    Me, Ke, _ = _element_forms(basis, max(p + 2, 8))
    Me_d = torch.from_numpy(np.ascontiguousarray(Me)).to(dev)
    Ke_d = torch.from_numpy(np.ascontiguousarray(Ke)).to(dev)
    m1 = np.zeros(nb)
    for e in range(N):
        m1[e : e + p1] += np.diag(Me[e])
    hbg = np.kron(m1, m1)
    h = 1.0 / N
    out = []
    k = 0
    for c, (full, cutm) in zip(cuts, classes):
        cls = np.where(full, 1, 0).astype(np.int8)
        slot = np.zeros((N, N), dtype=np.int32)
        cy, cx = np.nonzero(cutm)
        n_cut, min_frac = 0, 1.0
        for iy, ix in zip(cy.tolist(), cx.tolist()):
            a = area_h[k]
            cls[iy, ix] = 2 if a > 0.0 else 3
            slot[iy, ix] = k
            if a > 0.0:
                n_cut += 1
                min_frac = min(min_frac, a / (h * h))
            k += 1
        cls_d = torch.from_numpy(cls.ravel()).to(dev)
        slot_d = torch.from_numpy(slot.ravel()).to(dev)
        n = nb * nb
        cnt = torch.empty(n, dtype=torch.int32, device=dev)
        indptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
        hip._check(lib.cc_asm2d_count(N, p, hip._ptr(cls_d), hip._ptr(cnt), hip._ptr(indptr), hip._stream()),
                   "asm2d_count")
        nnz = int(indptr[-1].item())
        indices = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        Hv = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
        Kv = torch.empty_like(Hv)
        hip._check(lib.cc_asm2d_fill(N, p, hip._ptr(cls_d), hip._ptr(slot_d), hip._ptr(Me_d), hip._ptr(Ke_d),
                                     hip._ptr(Hb), hip._ptr(Kb), hip._ptr(indptr), hip._ptr(indices), hip._ptr(Hv),
                                     hip._ptr(Kv), hip._stream()), "asm2d_fill")
        ip = indptr.cpu().numpy()
        ix_ = indices.cpu().numpy()[:nnz]
        H = sp.csr_matrix((Hv.cpu().numpy()[:nnz], ix_, ip), shape=(n, n))
        K = sp.csr_matrix((Kv.cpu().numpy()[:nnz], ix_, ip), shape=(n, n))
        out.append(ObliqueCutSystem2D(basis_x=basis, basis_y=basis, cut=c, H=H, K=K, H_background_diag=hbg,
                                      n_cut_cells=n_cut, min_cut_fraction=min_frac))
    return out


def assemble_oblique_cut_gpu(basis_x, basis_y, cut: HalfPlaneCut, quadrature_order: int | None = None):
    if basis_x.n_elements != basis_y.n_elements or basis_x.degree != basis_y.degree:
        raise ValueError("equal element counts and degrees are expected in x and y")
    q = quadrature_order or max(10, basis_x.degree + 5)
    return assemble_batch_gpu(basis_x, [cut], q)[0]


def _system(sysm, degree, b_fn, meta, active_tol):
    agg = polynomial_preserving_extension(sysm, 0.05, degree, active_tol=active_tol)
    A = sym(pullback(sysm.K, agg.E) + pullback(sysm.H, agg.E))
    return CutSystem(A, b_fn(A), agg.E, agg.interior, agg.small_support, agg.active, meta)


def sliver_system_gpu(n_elements: int, degree: int, epsilon: float) -> CutSystem:
    """``cutcell.oblique2d.sliver_system`` with the device assembly."""
    basis = make_open_uniform_basis(n_elements, degree)
    cut = HalfPlaneCut.from_angle(0.0, basis.h * (1.0 - epsilon))
    sysm = assemble_batch_gpu(basis, [cut], max(24, 2 * degree + 14))[0]
    rng = np.random.default_rng(12345 + 17 * n_elements + degree)
    meta = dict(kind="sliver", n_elements=n_elements, degree=degree, epsilon=epsilon, n_full=sysm.n_full)
    return _system(sysm, degree, lambda A: rng.standard_normal(A.shape[0]), meta, 1.0e-100)


def oblique_batch_gpu(n_elements: int, degree: int, angles, offsets, seeds) -> list[CutSystem]:
    """Gaussian-right-hand-side oblique systems (``oblique_system(...,
    rhs="gaussian")``) for many cuts of one degree, assembled together."""
    basis = make_open_uniform_basis(n_elements, degree)
    cuts = [HalfPlaneCut.from_angle(a, o) for a, o in zip(angles, offsets)]
    systems = assemble_batch_gpu(basis, cuts, max(24, 2 * degree + 14))
    out = []
    for sysm, a, o, seed in zip(systems, angles, offsets, seeds):
        meta = dict(kind="oblique", n_elements=n_elements, degree=degree, angle=a, offset=o, n_full=sysm.n_full,
                    n_cut_cells=sysm.n_cut_cells, min_cut_fraction=sysm.min_cut_fraction)
        out.append(_system(sysm, degree, lambda A, s=seed: np.random.default_rng(s).standard_normal(A.shape[0]),
                           meta, 1.0e-22))
    return out


def c3_batch_gpu(n_elements: int = 48, n_cuts: int = 72, degrees=(3, 4, 5), n_slivers: int = 25, seed: int = 0):
    """``cutcell.workloads.c3_batch`` (same cases, same order) with every
    system's H and K assembled on the device: each degree's oblique cuts in
    one batch, the slivers one by one."""
    from .workloads import c3_cases

    cases = c3_cases(n_elements, n_cuts, degrees, n_slivers, seed)
    out = [None] * len(cases)
    for p in degrees:
        idx = [i for i, c in enumerate(cases) if c[0] == "oblique" and c[2] == p]
        if not idx:
            continue
        systems = oblique_batch_gpu(n_elements, p, [np.deg2rad(cases[i][3]) for i in idx],
                                    [cases[i][4] / n_elements for i in idx], [cases[i][5] for i in idx])
        for i, s in zip(idx, systems):
            s.meta.update(angle_deg=cases[i][3], offset_fraction=cases[i][4])
            out[i] = s
    for i, c in enumerate(cases):
        if c[0] == "sliver":
            out[i] = sliver_system_gpu(c[1], c[2], c[3])
    return out
