"""Float64 NumPy oracle of the masked 3x3x3 CNN -- TEST INFRASTRUCTURE ONLY.

Only the test suite imports this module. It restates, without any reference
counterpart (the reference has no learned preconditioner), the layer the HIP
kernels implement:

    out[v, co] = mask[v] * act( sum_t sum_ci W[t, ci, co] * in[v + off_t, ci] )

with taps ``t = (dz+1)*9 + (dy+1)*3 + (dx+1)`` over ``{-1, 0, 1}^3``, zero
padding on clamped axes and wrap-around on periodic axes, arrays indexed
``[z, y, x, channel]``. Backward passes are written out by hand and pinned by
central finite differences (``fd_check``).
"""

from __future__ import annotations

import numpy as np

OFFSETS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and back."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def shift(a: np.ndarray, dz: int, dy: int, dx: int, periodic=(False, False, False)) -> np.ndarray:
    """``out[z, y, x] = a[z + dz, y + dy, x + dx]`` (zero or wrapped outside)."""
    out = a
    for axis, (d, per) in enumerate(((dz, periodic[2]), (dy, periodic[1]), (dx, periodic[0]))):
        if d == 0:
            continue
        if per:
            out = np.roll(out, -d, axis=axis)
        else:
            res = np.zeros_like(out)
            n = out.shape[axis]
            src = [slice(None)] * out.ndim
            dst = [slice(None)] * out.ndim
            if d > 0:
                src[axis] = slice(d, n)
                dst[axis] = slice(0, n - d)
            else:
                src[axis] = slice(0, n + d)
                dst[axis] = slice(-d, n)
            res[tuple(dst)] = out[tuple(src)]
            out = res
    return out


def conv3(x, W, mask, relu=False, periodic=(False, False, False)):
    """Forward layer; returns ``(out, pre)`` (``pre`` = before act and mask)."""
    pre = np.zeros(x.shape[:3] + (W.shape[2],))
    for t, (dz, dy, dx) in enumerate(OFFSETS):
        pre += shift(x, dz, dy, dx, periodic) @ W[t]
    out = np.maximum(pre, 0.0) if relu else pre.copy()
    return out * mask[..., None], pre


def conv3_backward(x, W, mask, pre, gout, relu=False, periodic=(False, False, False)):
    """Gradients of ``sum(gout * out)`` w.r.t. ``x`` and ``W``."""
    g = gout * mask[..., None]
    if relu:
        g = g * (pre > 0)
    dW = np.zeros_like(W)
    dx_ = np.zeros_like(x)
    for t, (dz, dy, dx) in enumerate(OFFSETS):
        xs = shift(x, dz, dy, dx, periodic)
        dW[t] = np.einsum("zyxc,zyxo->co", xs, g)
        dx_ += shift(g @ W[t].T, -dz, -dy, -dx, periodic)
    return dx_, dW


def network(r, Ws, mask, periodic=(False, False, False)):
    """The preconditioner net: 3 hidden ReLU layers, linear output channel 0,
    plus the input (residual connection). ``r`` is ``[z, y, x]``."""
    x = np.zeros(r.shape + (16,))
    x[..., 0] = r * mask
    acts = [x]
    pres = []
    for i, W in enumerate(Ws):
        out, pre = conv3(acts[-1], W, mask, relu=i < len(Ws) - 1, periodic=periodic)
        acts.append(out)
        pres.append(pre)
    z = r * mask + acts[-1][..., 0]
    return z, acts, pres


def loss_and_grads(r, e, Ws, mask, periodic=(False, False, False)):
    """``L = ||z(r) - e||^2 / ||e||^2`` over masked voxels and ``dL/dW``."""
    z, acts, pres = network(r, Ws, mask, periodic)
    diff = (z - e) * mask
    den = float(np.sum((e * mask) ** 2))
    L = float(np.sum(diff**2)) / den
    g = np.zeros(acts[-1].shape)
    g[..., 0] = 2.0 * diff / den
    grads = [None] * len(Ws)
    for i in range(len(Ws) - 1, -1, -1):
        gx, dW = conv3_backward(acts[i], Ws[i], mask, pres[i], g, relu=i < len(Ws) - 1, periodic=periodic)
        grads[i] = dW
        g = gx
    return L, grads


def fd_check(r, e, Ws, mask, n_probe=6, h=1e-6, seed=0, periodic=(False, False, False)):
    """Max relative deviation of analytic from central-difference gradients."""
    rng = np.random.default_rng(seed)
    _, grads = loss_and_grads(r, e, Ws, mask, periodic)
    worst = 0.0
    for _ in range(n_probe):
        i = int(rng.integers(len(Ws)))
        idx = tuple(int(rng.integers(s)) for s in Ws[i].shape)
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[i][idx] += h
        Wm[i][idx] -= h
        fd = (loss_and_grads(r, e, Wp, mask, periodic)[0] - loss_and_grads(r, e, Wm, mask, periodic)[0]) / (2 * h)
        an = grads[i][idx]
        worst = max(worst, abs(fd - an) / max(abs(fd), abs(an), 1e-12))
    return worst


# ---------------------------------------------------------------------------
# The symmetric (tied-weight) preconditioner network.
#
#   z = D^-1 r + S C^T C S r,   S = D^-1/2,   C = L1 L0 (two masked linear
#   3x3x3 layers: W0 lifts channel 0 to 16 channels, W1 mixes 16 -> 16),
#
# run as four layers with weights (W0, W1, flip W1, flip W0): the adjoint of a
# correlation is the correlation with reversed taps and transposed channels.
# M = D^-1 + S C^T C S is symmetric positive definite with M >= D^-1.
# ---------------------------------------------------------------------------


def flip(W: np.ndarray) -> np.ndarray:
    """``flip(W)[t, a, b] = W[26 - t, b, a]`` (the adjoint layer's weights)."""
    return np.ascontiguousarray(W[::-1].transpose(0, 2, 1))


def sym_layers(W0: np.ndarray, W1: np.ndarray):
    return [W0, W1, flip(W1), flip(W0)]


def _inv_sqrt(d: np.ndarray) -> np.ndarray:
    return np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 1.0)


def sym_network(r, d, W0, W1, mask, periodic=(False, False, False)):
    """``z`` of the symmetric network for grids ``r, d`` (``[z, y, x]``, zero
    outside ``mask``); returns ``(z, acts)``."""
    s = _inv_sqrt(d) * mask
    x = np.zeros(r.shape + (16,))
    x[..., 0] = s * r
    acts = [x]
    for W in sym_layers(W0, W1):
        out, _ = conv3(acts[-1], W, mask, relu=False, periodic=periodic)
        acts.append(out)
    dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0) * mask
    z = dinv * r + s * acts[-1][..., 0]
    return z, acts


def sym_loss_and_grads(r, e, d, W0, W1, mask, periodic=(False, False, False)):
    """``L = ||z - e||^2 / ||e||^2`` (masked voxels) and ``(dL/dW0, dL/dW1)``,
    each the sum of its two tied layers' gradients."""
    z, acts = sym_network(r, d, W0, W1, mask, periodic)
    diff = (z - e) * mask
    den = float(np.sum((e * mask) ** 2))
    L = float(np.sum(diff**2)) / den
    s = _inv_sqrt(d) * mask
    g = np.zeros(acts[-1].shape)
    g[..., 0] = 2.0 * diff * s / den
    layers = sym_layers(W0, W1)
    dV = [None] * 4
    for i in range(3, -1, -1):
        gx, dW = conv3_backward(acts[i], layers[i], mask, None, g, relu=False, periodic=periodic)
        dV[i] = dW
        g = gx
    return L, (dV[0] + flip(dV[3]), dV[1] + flip(dV[2]))


def sym_fd_check(r, e, d, W0, W1, mask, n_probe=6, h=1e-6, seed=0, periodic=(False, False, False)):
    """Max relative deviation of the tied analytic gradients from central
    differences."""
    rng = np.random.default_rng(seed)
    _, grads = sym_loss_and_grads(r, e, d, W0, W1, mask, periodic)
    worst = 0.0
    for _ in range(n_probe):
        i = int(rng.integers(2))
        idx = tuple(int(rng.integers(sz)) for sz in (27, 16, 16))
        if i == 0:
            idx = (idx[0], 0, idx[2])  # W0 reads input channel 0 only
        Wp = [W0.copy(), W1.copy()]
        Wm = [W0.copy(), W1.copy()]
        Wp[i][idx] += h
        Wm[i][idx] -= h
        Lp = sym_loss_and_grads(r, e, d, *Wp, mask, periodic)[0]
        Lm = sym_loss_and_grads(r, e, d, *Wm, mask, periodic)[0]
        fd = (Lp - Lm) / (2 * h)
        an = grads[i][idx]
        worst = max(worst, abs(fd - an) / max(abs(fd), abs(an), 1e-12))
    return worst


# ---------------------------------------------------------------------------
# Multilevel form: grid hierarchy, prolongation, and the summed preconditioner.
# ---------------------------------------------------------------------------


def coarse_len(n: int, periodic: bool) -> int:
    """Points of the next level: every other point (periodic: n even)."""
    if periodic:
        assert n % 2 == 0
        return n // 2
    return (n + 1) // 2


def prolong_1d(n: int, periodic: bool) -> np.ndarray:
    """Dense ``[n, n_c]`` linear interpolation: fine 2c <- c (1), fine
    2c + 1 <- c, c + 1 (1/2 each; c + 1 wraps when periodic; past the end of
    a clamped axis, c alone with weight 1). Written from the definition,
    entry by entry."""
    nc = coarse_len(n, periodic)
    P = np.zeros((n, nc))
    for f in range(n):
        if f % 2 == 0:
            P[f, f // 2] = 1.0
        else:
            c = (f - 1) // 2
            if c + 1 < nc:
                P[f, c] += 0.5
                P[f, c + 1] += 0.5
            elif periodic:
                P[f, c] += 0.5
                P[f, 0] += 0.5
            else:
                P[f, c] += 1.0
    return P


def prolong_3d(dims, periodic):
    """``P`` on grids raveled ``[z, y, x]`` (dims = (nx, ny, nz)): kron(Pz, Py, Px)."""
    Px, Py, Pz = (prolong_1d(n, p) for n, p in zip(dims, periodic))
    return np.kron(Pz, np.kron(Py, Px))


def ml_network(r, d, levels, periodic=(False, False, False)):
    """``z = D^-1 r + S sum_l Q_l C_l^T C_l Q_l^T S r`` on the background grid.

    ``levels``: list over grids 0..L-1 of ``(dims, W0, W1, mask)`` or
    ``(dims, None, None, None)`` for a level without a network; ``r, d`` are
    ``[z, y, x]`` grids of level 0 (zero off the unknowns). The level
    network is the symmetric one of :func:`sym_network` with unit diagonal,
    i.e. ``C^T C`` applied to channel 0.
    """
    mask0 = levels[0][3] if levels[0][3] is not None else (d != 0).astype(float)
    s = _inv_sqrt(d) * (d != 0)
    u = [(s * r).ravel()]
    Ps = [None]
    for l in range(1, len(levels)):
        P = prolong_3d(levels[l - 1][0], periodic)
        Ps.append(P)
        u.append(P.T @ u[-1])
    t = np.zeros_like(u[0])
    for l in range(len(levels) - 1, -1, -1):
        dims, W0, W1, mask = levels[l]
        y = np.zeros(int(np.prod(dims)))
        if W0 is not None:
            shape = (dims[2], dims[1], dims[0])
            x = np.zeros(shape + (16,))
            x[..., 0] = u[l].reshape(shape)
            a = x
            for W in sym_layers(W0, W1):
                a, _ = conv3(a, W, mask, relu=False, periodic=periodic)
            y = a[..., 0].ravel()
        if l == len(levels) - 1:
            t = y
        else:
            t = Ps[l + 1] @ t + y
    dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    del mask0
    return dinv * r + s * t.reshape(r.shape)
