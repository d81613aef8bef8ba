"""Learned preconditioner: a masked 4-layer, 16-channel 3x3x3 CNN on the
background dof grid, trained on the GPU and applied inside flexible CG.

Everything numerical runs in the HIP kernels of ``csrc/cnn.hip`` and
``csrc/fcg.hip`` (bfloat16 implicit-GEMM convolutions on MFMA with fp32
accumulation, fp32 weight gradients, Adam); PyTorch only owns the device
buffers. The network is symmetric by construction:

    u  = S r               (S = D^-1/2, channel 0 at the unknowns' voxels)
    h1 = mask * conv(u, W0),       h2 = mask * conv(h1, W1)
    h3 = mask * conv(h2, flip W1), y  = mask * conv(h3, flip W0)
    z  = D^-1 r + S y[channel 0]

i.e. ``z = (D^-1 + S C^T C S) r`` with ``C = L1 L0`` (the flipped layers are
the adjoints: reversed taps, transposed channels). In exact arithmetic the
preconditioner is symmetric positive definite and never below Jacobi
(``M >= D^-1``), so training cannot produce an indefinite map; the
bfloat16 evaluation makes the applied map slightly inexact, which the
flexible CG absorbs. It is trained on (residual, error) pairs taken from the
solver's own Jacobi-PCG iterates, minimising ``||z(r) - e||^2 / ||e||^2``
per pair. Inside flexible CG the float64 true residual decides success; a
learned solve that does not converge falls back to Jacobi-PCG.

The reference has no learned preconditioner; its Jacobi rule is
``code/fast_solver_backends.py:74-87`` (CPU) and ``:237-238`` (GPU).
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import hip

NLAYERS = 4
WSIZE = 27 * 16 * 16


class CnnDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("weights", ctypes.c_void_p * 4),
        ("mask", ctypes.c_void_p),
        ("red2full", ctypes.c_void_p),
        ("n_red", ctypes.c_int32),
    ]


ML_MAX_LEVELS = 8


class MlDesc(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int32),
        ("has_net", ctypes.c_int32 * ML_MAX_LEVELS),
        ("net", CnnDesc * ML_MAX_LEVELS),
    ]


def _bind(lib):
    P, I, F, D, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
    sigs = {
        "cc_mg_restrict": [P, P, P, P, I, P, I, P],
        "cc_mg_prolong": [P, P, P, P, I, P, I, P, P],
        "cc_mg_to_ch0": [ctypes.c_int64, P, I, P, P],
        "cc_mg_train_input": [I, P, P, P, F, P, P],
        "cc_mg_train_loss": [I, P, P, P, P, F, P, F, F, P, P, P, P],
        "cc_mlcnn_precondition": [P, P, P, P, P, SZ, P],
        "cc_fcg_sumfact_ml": [P, P, P, P, D, I, I, P, SZ, P, P, P],
        "cc_fcg_sumfact_ml_k": [P, P, I, P, P, D, I, I, P, SZ, P, P],
        "cc_fcg_sumfact_ml_mixed_k": [P, P, P, I, P, P, D, I, D, I, I, P, SZ, P, P],
        "cc_mfma_probe": [P, P, P, P],
        "cc_mfma32_probe": [P, P, P, P],
        "cc_net_set_config": [I, I],
        "cc_net_level_forward": [P, I, P, P, P, SZ, P],
        "cc_cnn_set_config": [I],
        "cc_cnn_conv3_forward": [I, I, I, P, P, P, P, P, I, P, P, P],
        "cc_cnn_flip_weights": [P, P, P],
        "cc_cnn_sym_weights": [P, P, P],
        "cc_cnn_sym_grads": [P, P, I, P],
        "cc_cnn_conv3_wgrad": [I, I, I, P, P, P, P, F, I, I, P, SZ, P],
        "cc_cnn_adam": [I, P, P, P, P, F, F, F, F, I, P, P, P],
        "cc_cnn_train_input": [I, P, P, P, F, P, P],
        "cc_cnn_train_loss": [I, P, P, P, P, F, P, F, F, P, P, P, P],
        "cc_cnn_precondition": [P, P, P, P, P, SZ, P],
        "cc_fcg_sumfact": [P, P, P, P, D, I, I, P, SZ, P, P, P],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    for name, args in {
        "cc_cnn_wgrad_workspace_bytes": [I, I, I, I],
        "cc_cnn_precondition_workspace_bytes": [P],
        "cc_fcg_sumfact_workspace_bytes": [P, P],
        "cc_mlcnn_precondition_workspace_bytes": [P],
        "cc_fcg_sumfact_ml_workspace_bytes": [P, P],
        "cc_fcg_sumfact_ml_k_workspace_bytes": [P, P, I],
        "cc_fcg_sumfact_ml_mixed_k_workspace_bytes": [P, P, I],
        "cc_net_level_workspace_bytes": [P, I],
    }.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = SZ
    return lib


def _lib():
    return _bind(hip.require_gpu())


def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even); returns uint16 bits."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


@dataclass
class TrainReport:
    steps: int
    batch: int
    loss_first: float
    loss_last: float
    seconds: float
    history: list


class MaskedCNN:
    """Master weights, optimiser state and device buffers of the symmetric
    preconditioner network (two fp32 masters W0, W1; four bf16 layers)."""

    def __init__(self, grid_shape, red2full: np.ndarray | None, periodic=(False, False, False), seed: int = 0,
                 device: str = "cuda", init_scale: float = 0.3, mask: np.ndarray | None = None):
        import torch

        self.torch = torch
        self.lib = _lib()
        self.nx, self.ny, self.nz = (int(v) for v in grid_shape)
        self.voxels = self.nx * self.ny * self.nz
        self.periodic = (ctypes.c_int32 * 3)(*[int(bool(p)) for p in periodic])
        self.device = device
        self.n = int(red2full.size) if red2full is not None else 0
        rng = np.random.default_rng(seed)
        w0 = rng.standard_normal((27, 16, 16)) * np.sqrt(1.0 / 27.0) * init_scale
        w0[:, 1:, :] = 0.0  # the input is channel 0 only
        w1 = rng.standard_normal((27, 16, 16)) * np.sqrt(1.0 / (27.0 * 16.0))
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
        self.w32 = t(np.stack([w0.ravel(), w1.ravel()]).astype(np.float32), np.float32)  # (2, WSIZE)
        self.wbf = torch.zeros((NLAYERS, WSIZE), dtype=torch.uint16, device=device)
        self.m1 = torch.zeros_like(self.w32)
        self.m2 = torch.zeros_like(self.w32)
        self.grad = torch.zeros_like(self.w32)  # masters' gradient
        self.lgrad = torch.zeros((NLAYERS, WSIZE), dtype=torch.float32, device=device)  # per layer
        self.step_count = 0
        if mask is None:
            mask = np.zeros(self.voxels, dtype=np.uint8)
            mask[red2full] = 1
        self.mask = t(np.asarray(mask, dtype=np.uint8).ravel(), np.uint8)
        self.red2full = t(red2full if red2full is not None else np.zeros(1, np.int64), np.int32)
        # Activation buffers (bf16 [voxel][16]) and outputs.
        z16 = lambda: torch.zeros(self.voxels * 16, dtype=torch.uint16, device=device)
        self.acts = [z16() for _ in range(NLAYERS)]
        self.y = torch.zeros(self.voxels * 16, dtype=torch.float32, device=device)
        self.g = [z16() for _ in range(2)]
        self.gout = z16()
        self.wgrad_blocks = max(1, min(512, ((self.nx + 15) // 16) * self.ny * self.nz))
        nb = int(self.lib.cc_cnn_wgrad_workspace_bytes(self.nx, self.ny, self.nz, self.wgrad_blocks))
        self.wgrad_ws = torch.empty(nb, dtype=torch.uint8, device=device)
        self.part = torch.zeros(2 * 1024, dtype=torch.float64, device=device)
        self._desc = None
        self.refresh_layers()

    def refresh_layers(self):
        """bf16 layers (W0, W1, flip W1, flip W0) from the fp32 masters."""
        hip._check(self.lib.cc_cnn_sym_weights(hip._ptr(self.w32), hip._ptr(self.wbf), hip._stream()), "sym_weights")

    def set_masters(self, w0: np.ndarray, w1: np.ndarray):
        torch = self.torch
        self.w32.copy_(torch.from_numpy(np.stack([np.ravel(w0), np.ravel(w1)]).astype(np.float32)))
        self.refresh_layers()

    def layer_weights(self) -> list:
        """The four bf16 layers as float64 arrays [27][16][16] (host copies)."""
        return [from_bf16_bits(self.wbf[l].cpu().numpy()).astype(np.float64).reshape(27, 16, 16)
                for l in range(NLAYERS)]

    # -- descriptor for the fused preconditioner / FCG ------------------------
    def desc(self) -> CnnDesc:
        d = CnnDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        for k in range(3):
            d.periodic[k] = self.periodic[k]
        for l in range(NLAYERS):
            d.weights[l] = self.wbf[l].data_ptr()
        d.mask = self.mask.data_ptr()
        d.red2full = self.red2full.data_ptr() if self.n else None
        d.n_red = self.n
        self._desc = d
        return d

    # -- layers ---------------------------------------------------------------
    def _conv(self, x, layer_w, out=None, out_f32=None):
        p = hip._ptr
        hip._check(self.lib.cc_cnn_conv3_forward(self.nx, self.ny, self.nz, self.periodic, p(x), p(layer_w),
                                                 p(self.mask), None, 0, p(out) if out is not None else None,
                                                 p(out_f32) if out_f32 is not None else None, hip._stream()),
                   "conv3_forward")

    def forward(self):
        """acts[0] holds the input; fills acts[1..3] and y (fp32)."""
        for l in range(NLAYERS - 1):
            self._conv(self.acts[l], self.wbf[l], out=self.acts[l + 1])
        self._conv(self.acts[NLAYERS - 1], self.wbf[NLAYERS - 1], out_f32=self.y)

    def backward(self, scale: float, accumulate: bool):
        """gout holds dL/dy (channel 0); adds scale * dL/dW into the masters'
        gradient (or sets it when ``accumulate`` is false). The data gradient
        of layer l is the layer with flipped weights, which is layer 3 - l."""
        p = hip._ptr
        g = self.gout
        for l in range(NLAYERS - 1, -1, -1):
            hip._check(self.lib.cc_cnn_conv3_wgrad(self.nx, self.ny, self.nz, self.periodic, p(self.acts[l]), p(g),
                                                   p(self.lgrad[l]), float(scale), 0, self.wgrad_blocks,
                                                   p(self.wgrad_ws), self.wgrad_ws.numel(), hip._stream()),
                       "conv3_wgrad")
            if l == 0:
                break
            gnext = self.g[l % 2]
            self._conv(g, self.wbf[NLAYERS - 1 - l], out=gnext)
            g = gnext
        hip._check(self.lib.cc_cnn_sym_grads(p(self.lgrad), p(self.grad), int(accumulate), hip._stream()), "sym_grads")

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step_count += 1
        p = hip._ptr
        hip._check(self.lib.cc_cnn_adam(2 * WSIZE, p(self.w32), p(self.grad), p(self.m1), p(self.m2), lr, beta1,
                                        beta2, eps, self.step_count, None, p(self.wbf), hip._stream()), "adam")
        self.refresh_layers()  # Adam rounded W0, W1 into layers 0, 1; rebuild all four

    # -- training -------------------------------------------------------------
    def sample_loss_grad(self, r32, e32, diag, scale: float, gscale: float, accumulate: bool,
                         e_norm2: float) -> float:
        """Forward + backward for one (r, e) pair (``e_norm2 = ||e||^2``);
        returns its loss as a device scalar (no synchronisation)."""
        p = hip._ptr
        hip._check(self.lib.cc_cnn_train_input(self.n, p(self.red2full), p(r32), p(diag), float(scale), p(self.acts[0]),
                                               hip._stream()), "train_input")
        self.forward()
        t2 = scale * scale * e_norm2
        nparts = ctypes.c_int32()
        hip._check(self.lib.cc_cnn_train_loss(self.n, p(self.red2full), p(r32), p(e32), p(diag), float(scale), p(self.y),
                                              float(1.0 / t2), float(gscale), p(self.gout), p(self.part),
                                              ctypes.addressof(nparts), hip._stream()), "train_loss")
        self.backward(1.0, accumulate)
        parts = self.part[: 2 * nparts.value].view(-1, 2).sum(dim=0)
        return parts[0] / parts[1]

    def train(self, R, Eerr, diag, *, steps: int = 200, batch: int = 8, lr: float = 2e-3, seed: int = 0,
              log_every: int = 0) -> TrainReport:
        """Adam on mini-batches of (residual, error) pairs (float32, [S, n])."""
        torch = self.torch
        rng = np.random.default_rng(seed)
        S = R.shape[0]
        scales, enorm2 = self._sample_scales(R, Eerr, diag)
        hist = []
        t0 = time.perf_counter()
        for step in range(steps):
            idx = rng.choice(S, size=min(batch, S), replace=False)
            losses = []
            for k, i in enumerate(idx):
                losses.append(self.sample_loss_grad(R[i], Eerr[i], diag, float(scales[i]), 1.0 / len(idx), k > 0,
                                                    float(enorm2[i])))
            self.adam_step(lr)
            hist.append(float(torch.stack(losses).mean().item()))
            if log_every and step % log_every == 0:
                print(f"[cnn] step {step} loss {hist[-1]:.4f}", flush=True)
        torch.cuda.synchronize()
        return TrainReport(steps=steps, batch=batch, loss_first=hist[0], loss_last=hist[-1],
                           seconds=time.perf_counter() - t0, history=hist)

    @staticmethod
    def _sample_scales(R, Eerr, diag):
        """Per pair: 1/||S r|| (the input is scaled to unit norm) and ||e||^2
        (bookkeeping of the data set)."""
        s = diag.double().clamp_min(0.0)
        s = torch_where_pos_rsqrt(s)
        in_norm = (R.double() * s).norm(dim=1).cpu().numpy()
        enorm2 = Eerr.double().pow(2).sum(dim=1).cpu().numpy()
        return 1.0 / np.maximum(in_norm, 1e-300), enorm2

    def evaluate(self, R, Eerr, diag) -> float:
        """Mean loss over the given pairs (no weight update)."""
        scales, enorm2 = self._sample_scales(R, Eerr, diag)
        keep = self.grad.clone()
        out = []
        for i in range(R.shape[0]):
            out.append(self.sample_loss_grad(R[i], Eerr[i], diag, float(scales[i]), 0.0, True, float(enorm2[i])))
        self.grad.copy_(keep)
        return float(self.torch.stack(out).mean().item())

    # -- application ----------------------------------------------------------
    def precondition(self, r, diag):
        torch = self.torch
        d = self.desc()
        nb = int(self.lib.cc_cnn_precondition_workspace_bytes(ctypes.addressof(d)))
        ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        z = torch.empty_like(r)
        hip._check(self.lib.cc_cnn_precondition(ctypes.addressof(d), hip._ptr(r), hip._ptr(diag), hip._ptr(z),
                                                hip._ptr(ws), nb, hip._stream()), "cnn_precondition")
        return z

    def fcg(self, dmf, b, *, tol=1.0e-8, maxiter=5000, check_every=16, history=False, x=None):
        """Flexible CG with this network on a ``DeviceMatrixFree`` operator."""
        torch = self.torch
        d = self.desc()
        nb = int(self.lib.cc_fcg_sumfact_workspace_bytes(ctypes.addressof(dmf.desc), ctypes.addressof(d)))
        if nb == 0:
            raise ValueError("invalid descriptors")
        ws = getattr(self, "_fcg_ws", None)
        if ws is None or ws.numel() < nb:
            ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            self._fcg_ws = ws
        x = torch.empty_like(b) if x is None else x
        res = hip.PcgResult()
        hist = torch.full((int(maxiter) + 1,), float("nan"), dtype=torch.float64, device=b.device) if history else None
        hip._check(self.lib.cc_fcg_sumfact(ctypes.addressof(dmf.desc), ctypes.addressof(d), hip._ptr(b), hip._ptr(x),
                                           float(tol), int(maxiter), int(check_every), hip._ptr(ws), nb,
                                           ctypes.addressof(res), hip._ptr(hist) if hist is not None else None,
                                           hip._stream()), "fcg_sumfact")
        if history:
            return x, res, hist.cpu().numpy()
        return x, res


def coarse_dims(dims, periodic):
    """Grid of the next level: every other point per axis (periodic axes
    must have an even count)."""
    out = []
    for n, per in zip(dims, periodic):
        if per:
            if n % 2:
                raise ValueError("a periodic axis needs an even point count to coarsen")
            out.append(n // 2)
        else:
            out.append((n + 1) // 2)
    return tuple(out)


def _restrict_1d(n: int, per: bool):
    """Sparse ``R = P^T`` of one axis (the rule of ``csrc/mg.hip``)."""
    import scipy.sparse as sp

    nc = n // 2 if per else (n + 1) // 2
    rows, cols, vals = [], [], []
    for f in range(n):
        if f % 2 == 0:
            srcs = [(f // 2, 1.0)]
        else:
            c = (f - 1) // 2
            if c + 1 < nc:
                srcs = [(c, 0.5), (c + 1, 0.5)]
            elif per:
                srcs = [(c, 0.5), (0, 0.5)]
            else:
                srcs = [(c, 1.0)]
        for c, w in srcs:
            rows.append(c)
            cols.append(f)
            vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nc, n))


def restrict_host(a: np.ndarray, dims, periodic) -> np.ndarray:
    """``R a`` for a grid raveled ``[z, y, x]`` of dims ``(nx, ny, nz)``."""
    nx, ny, nz = dims
    g = np.asarray(a, dtype=np.float64).reshape(nz, ny, nx)
    Rx, Ry, Rz = (_restrict_1d(n, bool(p)) for n, p in zip(dims, periodic))
    g = (Rx @ g.reshape(-1, nx).T).T.reshape(nz, ny, -1)
    g = np.einsum("cy,zyx->zcx", Ry.toarray(), g)
    g = np.einsum("cz,zyx->cyx", Rz.toarray(), g)
    return g.ravel()


class MultilevelCNN:
    """The multilevel learned preconditioner

        z = D^-1 r + S sum_l Q_l C_l^T C_l Q_l^T S r,

    one symmetric masked network (``MaskedCNN``) per level of the background
    grid hierarchy (level l + 1 keeps every other point of level l; Q_l
    prolongates level l to the background grid, Q_l^T restricts). Each level's
    mask holds the voxels that the restriction of the unknowns reaches. The
    coarse levels see the smooth error that a 3x3x3 stack on the fine grid
    cannot; the whole map stays symmetric and never below Jacobi.

    ``net_levels``: the levels that carry a network (default: all). The
    background grid's network does most of the work on these B-spline
    operators, whose slow modes under Jacobi are high-frequency coefficient
    patterns of the p = 3 basis and of the cut cells, not smooth fields; the
    coarse levels add the smooth part that matters for pure (unscreened)
    Poisson operators such as the pressure projection.
    """

    def __init__(self, grid_shape, red2full: np.ndarray, periodic=(False, False, False), levels: int = 5,
                 net_levels=None, seed: int = 0, device: str = "cuda", init_scale: float = 0.3):
        import torch

        self.torch = torch
        self.lib = _lib()
        self.device = device
        self.periodic = tuple(bool(p) for p in periodic)
        self.red2full_np = np.asarray(red2full, dtype=np.int64)
        self.n = int(self.red2full_np.size)
        dims = [tuple(int(v) for v in grid_shape)]
        for _ in range(1, levels):
            try:
                nxt = coarse_dims(dims[-1], self.periodic)
            except ValueError:
                break
            if min(nxt) < 2:
                break
            dims.append(nxt)
        self.dims = dims
        self.levels = len(dims)
        if net_levels is None:
            net_levels = list(range(self.levels))
        self.net_levels = sorted(int(l) for l in net_levels if l < self.levels)
        if not self.net_levels or self.net_levels[-1] != self.levels - 1:
            raise ValueError("the coarsest level must carry a network")
        masks = [np.zeros(int(np.prod(dims[0])), dtype=np.float64)]
        masks[0][self.red2full_np] = 1.0
        for l in range(1, self.levels):
            masks.append((restrict_host(masks[-1], dims[l - 1], self.periodic) > 0).astype(np.float64))
        self.masks = [m.astype(np.uint8) for m in masks]
        self.nets = {}
        # The restriction sums (1/2 + 1 + 1/2)^3 = 8 fine values per coarse
        # point, so a smooth input reaches level l amplified by 8^l; starting
        # level l's first layer 8^(l/2) smaller keeps every level's initial
        # correction the size of the fine one (training starts near Jacobi).
        for l in self.net_levels:
            self.nets[l] = MaskedCNN(dims[l], self.red2full_np if l == 0 else None, self.periodic, seed=seed + 17 * l,
                                     device=device, init_scale=init_scale * 8.0 ** (-0.5 * l), mask=self.masks[l])
        f32 = lambda nvox: torch.zeros(nvox, dtype=torch.float32, device=device)
        self.u = [f32(int(np.prod(d))) for d in dims]
        self.t = [f32(int(np.prod(d))) for d in dims]
        self.g = [f32(int(np.prod(d))) for d in dims]
        self.red2full = torch.from_numpy(self.red2full_np.astype(np.int32)).to(device)
        self.part = torch.zeros(2 * 1024, dtype=torch.float64, device=device)
        self._cdims = [(ctypes.c_int32 * 3)(*d) for d in dims]
        self._cper = (ctypes.c_int32 * 3)(*[int(p) for p in self.periodic])
        self._desc = None

    # -- descriptor -------------------------------------------------------------
    def desc(self) -> MlDesc:
        d = MlDesc()
        d.levels = self.levels
        for l in range(self.levels):
            nx, ny, nz = self.dims[l]
            if l in self.nets:
                d.net[l] = self.nets[l].desc()
                d.has_net[l] = 1
            else:
                d.net[l].nx, d.net[l].ny, d.net[l].nz = nx, ny, nz
                for k in range(3):
                    d.net[l].periodic[k] = int(self.periodic[k])
                d.has_net[l] = 0
        d.net[0].red2full = self.red2full.data_ptr()
        d.net[0].n_red = self.n
        self._desc = d
        return d

    # -- hierarchy transfers ------------------------------------------------------
    def _restrict(self, l, fine, fs, coarse, cs):
        hip._check(self.lib.cc_mg_restrict(self._cdims[l - 1], self._cdims[l], self._cper, hip._ptr(fine), fs,
                                           hip._ptr(coarse), cs, hip._stream()), "mg_restrict")

    def _prolong(self, l, coarse, cs, add, as_, fine):
        hip._check(self.lib.cc_mg_prolong(self._cdims[l - 1], self._cdims[l], self._cper, hip._ptr(coarse), cs,
                                          hip._ptr(add) if add is not None else None, as_, hip._ptr(fine),
                                          hip._stream()), "mg_prolong")

    def _to_ch0(self, l, src, grid):
        hip._check(self.lib.cc_mg_to_ch0(int(np.prod(self.dims[l])), hip._ptr(src), 1, hip._ptr(grid),
                                         hip._stream()), "mg_to_ch0")

    def _forward_levels(self):
        """u[0] holds the fine input; fills every level's network output and
        returns (tensor, stride) of the summed correction on the fine grid."""
        for l in range(1, self.levels):
            self._restrict(l, self.u[l - 1], 1, self.u[l], 1)
        for l, net in self.nets.items():
            self._to_ch0(l, self.u[l], net.acts[0])
            net.forward()
        L = self.levels
        src, ss = self.nets[L - 1].y, 16
        for l in range(L - 1, 0, -1):
            add = self.nets[l - 1].y if (l - 1) in self.nets else None
            self._prolong(l, src, ss, add, 16, self.t[l - 1])
            src, ss = self.t[l - 1], 1
        return src, ss

    # -- training -----------------------------------------------------------------
    def sample_loss_grad(self, r32, e32, diag, scale: float, gscale: float, accumulate: bool,
                         e_norm2: float):
        """Forward + backward for one (r, e) pair; returns its loss (device scalar)."""
        p = hip._ptr
        s = hip._stream()
        hip._check(self.lib.cc_mg_train_input(self.n, p(self.red2full), p(r32), p(diag), float(scale), p(self.u[0]),
                                              s), "mg_train_input")
        t0, ts = self._forward_levels()
        if ts != 1:  # a single level: channel 0 of its output
            self.t[0].copy_(t0.view(-1, 16)[:, 0])
            t0 = self.t[0]
        t2 = scale * scale * e_norm2
        nparts = ctypes.c_int32()
        hip._check(self.lib.cc_mg_train_loss(self.n, p(self.red2full), p(r32), p(e32), p(diag), float(scale), p(t0),
                                             float(1.0 / t2), float(gscale), p(self.g[0]), p(self.part),
                                             ctypes.addressof(nparts), s), "mg_train_loss")
        for l in range(1, self.levels):
            self._restrict(l, self.g[l - 1], 1, self.g[l], 1)
        for l, net in self.nets.items():
            self._to_ch0(l, self.g[l], net.gout)
            net.backward(1.0, accumulate)
        parts = self.part[: 2 * nparts.value].view(-1, 2).sum(dim=0)
        return parts[0] / parts[1]

    def adam_step(self, lr: float):
        for net in self.nets.values():
            net.adam_step(lr)

    def train(self, R, Eerr, diag, *, steps: int = 200, batch: int = 8, lr: float = 2e-3, seed: int = 0,
              log_every: int = 0) -> TrainReport:
        """Adam on mini-batches of (residual, error) pairs (float32, [S, n])."""
        torch = self.torch
        rng = np.random.default_rng(seed)
        S = R.shape[0]
        scales, enorm2 = MaskedCNN._sample_scales(R, Eerr, diag)
        # The fine input buffer is written only at the unknowns' voxels.
        self.u[0].zero_()
        self.g[0].zero_()
        hist = []
        t0 = time.perf_counter()
        for step in range(steps):
            idx = rng.choice(S, size=min(batch, S), replace=False)
            losses = [self.sample_loss_grad(R[i], Eerr[i], diag, float(scales[i]), 1.0 / len(idx), k > 0,
                                            float(enorm2[i])) for k, i in enumerate(idx)]
            self.adam_step(lr)
            hist.append(float(torch.stack(losses).mean().item()))
            if log_every and step % log_every == 0:
                print(f"[mlcnn] step {step} loss {hist[-1]:.4f}", flush=True)
        torch.cuda.synchronize()
        return TrainReport(steps=steps, batch=batch, loss_first=hist[0], loss_last=hist[-1],
                           seconds=time.perf_counter() - t0, history=hist)

    # -- application ----------------------------------------------------------------
    def precondition(self, r, diag):
        torch = self.torch
        d = self.desc()
        nb = int(self.lib.cc_mlcnn_precondition_workspace_bytes(ctypes.addressof(d)))
        if nb == 0:
            raise ValueError("invalid multilevel descriptor")
        ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        z = torch.empty_like(r)
        hip._check(self.lib.cc_mlcnn_precondition(ctypes.addressof(d), hip._ptr(r), hip._ptr(diag), hip._ptr(z),
                                                  hip._ptr(ws), nb, hip._stream()), "mlcnn_precondition")
        return z

    def fcg(self, dmf, b, *, tol=1.0e-8, maxiter=5000, check_every=16, history=False, x=None):
        """Flexible CG with this preconditioner on a ``DeviceMatrixFree`` operator."""
        torch = self.torch
        d = self.desc()
        nb = int(self.lib.cc_fcg_sumfact_ml_workspace_bytes(ctypes.addressof(dmf.desc), ctypes.addressof(d)))
        if nb == 0:
            raise ValueError("invalid descriptors")
        ws = getattr(self, "_fcg_ws", None)
        if ws is None or ws.numel() < nb:
            ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            self._fcg_ws = ws
        x = torch.empty_like(b) if x is None else x
        res = hip.PcgResult()
        hist = torch.full((int(maxiter) + 1,), float("nan"), dtype=torch.float64, device=b.device) if history else None
        hip._check(self.lib.cc_fcg_sumfact_ml(ctypes.addressof(dmf.desc), ctypes.addressof(d), hip._ptr(b),
                                              hip._ptr(x), float(tol), int(maxiter), int(check_every), hip._ptr(ws),
                                              nb, ctypes.addressof(res),
                                              hip._ptr(hist) if hist is not None else None, hip._stream()),
                   "fcg_sumfact_ml")
        if history:
            return x, res, hist.cpu().numpy()
        return x, res

    def fcg_batched(self, dmf, B, *, tol=1.0e-8, maxiter=5000, check_every=16, X=None):
        """Flexible CG with this preconditioner on ``k`` right-hand sides at
        once (``B`` row-major ``[n, k]``, k in 1, 2, 4, 8, 16): one K-column
        operator apply and one batched network pass per iteration; each column
        stops on its own. Returns ``(X, results)``."""
        torch = self.torch
        d = self.desc()
        B = B.contiguous()
        k = B.shape[1]
        nb = int(self.lib.cc_fcg_sumfact_ml_k_workspace_bytes(ctypes.addressof(dmf.desc), ctypes.addressof(d), k))
        if nb == 0:
            raise ValueError("invalid descriptors or k")
        ws = getattr(self, "_fcgk_ws", None)
        if ws is None or ws.numel() < nb:
            ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            self._fcgk_ws = ws
        X = torch.empty_like(B) if X is None else X
        res = (hip.PcgResult * k)()
        hip._check(self.lib.cc_fcg_sumfact_ml_k(ctypes.addressof(dmf.desc), ctypes.addressof(d), k, hip._ptr(B),
                                                hip._ptr(X), float(tol), int(maxiter), int(check_every),
                                                hip._ptr(ws), nb, ctypes.addressof(res), hip._stream()),
                   "fcg_sumfact_ml_k")
        return X, list(res)


    def fcg_mixed_batched(self, dmf, B, *, tol=1.0e-8, maxiter=5000, inner_tol=1.0e-4, max_outer=12,
                          check_every=16, X=None):
        """``fcg_batched`` inside float64 iterative refinement: each sweep
        solves for the scaled residual to ``inner_tol`` with the operator
        applied in float32, and the sweeps stop on the float64 residual
        ``||B - A X|| / ||B|| <= tol``. Returns ``(X, results)``;
        ``iterations`` counts inner iterations and ``outer`` the sweeps."""
        torch = self.torch
        d = self.desc()
        B = B.contiguous()
        k = B.shape[1]
        nb = int(self.lib.cc_fcg_sumfact_ml_mixed_k_workspace_bytes(ctypes.addressof(dmf.desc), ctypes.addressof(d),
                                                                     k))
        if nb == 0:
            raise ValueError("invalid descriptors or k")
        ws = getattr(self, "_fcgk_mixed_ws", None)
        if ws is None or ws.numel() < nb:
            ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            self._fcgk_mixed_ws = ws
        X = torch.empty_like(B) if X is None else X
        res = (hip.PcgResult * k)()
        hip._check(self.lib.cc_fcg_sumfact_ml_mixed_k(ctypes.addressof(dmf.desc), dmf.blocks32, ctypes.addressof(d),
                                                      k, hip._ptr(B), hip._ptr(X), float(tol), int(maxiter),
                                                      float(inner_tol), int(max_outer), int(check_every),
                                                      hip._ptr(ws), nb, ctypes.addressof(res), hip._stream()),
                   "fcg_sumfact_ml_mixed_k")
        return X, list(res)


def torch_where_pos_rsqrt(d):
    """1/sqrt(d) where d > 0, else 1 (the kernels' symmetric scaling rule)."""
    import torch

    return torch.where(d > 0, torch.rsqrt(torch.where(d > 0, d, torch.ones_like(d))), torch.ones_like(d))


def generate_pairs(dmf, rhs, snapshots, tol_ref: float = 1.0e-11):
    """(residual, error) pairs from the solver's own Jacobi-PCG iterates.

    For every right-hand side b (rows of ``rhs``) one tight Jacobi-PCG solve
    gives x* and, on the way, the iterates x_k after each k in ``snapshots``
    (copied out by the device loop); r_k = b - A x_k (float64 apply) and
    e_k = x* - x_k. Snapshots at or past the solve's own stop are dropped.
    Returns float32 tensors ``R, E`` of shape [S, n].
    """
    import torch

    ks = sorted(int(k) for k in snapshots)
    Rs, Es = [], []
    for b in rhs:
        xs, res, snaps = dmf.pcg(b, tol=tol_ref, maxiter=40000, snapshots=ks)
        for j, k in enumerate(ks):
            if k >= res.iterations:
                continue
            xk = snaps[j]
            Rs.append((b - dmf.apply(xk)).float())
            Es.append((xs - xk).float())
    return torch.stack(Rs), torch.stack(Es)


def generate_pairs_fcg(net, dmf, rhs, iters, tol_ref: float = 1.0e-11):
    """(residual, error) pairs from the learned solver's own iterates: for
    every right-hand side b, x* from one tight Jacobi-PCG solve and x_k from
    flexible CG with ``net`` stopped after k iterations, for each k in
    ``iters``; r_k = b - A x_k, e_k = x* - x_k. Fine-tuning on these pairs
    (with the Jacobi-PCG pairs) targets the modes the trained preconditioner
    still leaves slow. Returns float32 tensors ``R, E`` of shape [S, n]."""
    import torch

    Rs, Es = [], []
    for b in rhs:
        xs, _ = dmf.pcg(b, tol=tol_ref, maxiter=40000)
        for k in iters:
            xk, _ = net.fcg(dmf, b, tol=0.0, maxiter=int(k))
            Rs.append((b - dmf.apply(xk)).float())
            Es.append((xs - xk).float())
    return torch.stack(Rs), torch.stack(Es)


def solve_learned_with_fallback(net, dmf, b, *, tol=1.0e-8, maxiter=5000):
    """Flexible CG with the network (``MaskedCNN`` or ``MultilevelCNN``);
    Jacobi-PCG when it does not converge."""
    x, res = net.fcg(dmf, b, tol=tol, maxiter=maxiter)
    ok = res.status == hip.CC_CONVERGED and res.true_rel <= max(10.0 * tol, 1e-14)
    if ok:
        return x, res, "learned-fcg"
    x, res = dmf.pcg(b, tol=tol, maxiter=maxiter)
    return x, res, "jacobi-pcg-fallback"
