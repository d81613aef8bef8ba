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
