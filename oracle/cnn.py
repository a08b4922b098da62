"""Float64 restatement of the warm-start network (test infrastructure only).

The learned warm start has no counterpart in the reference: it seeds the
TVME PDHG loop (``tvme.py:496-575``, which starts from zeros at
``tvme.py:180-186``) with a 3-D CNN's estimate of the converged image.  The
network is five 3x3x3 'same' convolutions (zero padding), channels
2 -> 32 -> 32 -> 32 -> 32 -> 2 with ReLU between them and a residual output
pred = x + f(x); real / imaginary parts are the two channels of one
(bin, echo) image.

Volumes are channels-last, ``(Z, Y, X, C)``.  Taps are numbered
t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1) for offsets dz, dy, dx in {-1, 0, 1},
weights are ``w[t, ci, co]`` and a layer computes the cross-correlation

    y[z, y, x, co] = b[co] + sum_t sum_ci x[z + dz, y + dy, x + dx, ci] w[t, ci, co]

(torch's ``conv3d`` with ``weight[co, ci, dz + 1, dy + 1, dx + 1]``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.
"""

from __future__ import annotations

import numpy as np

OFFSETS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
CHANNELS = (2, 32, 32, 32, 32, 2)


def _shifted(x, dz, dy, dx):
    """s[z, y, x] = x[z + dz, y + dy, x + dx] with zeros outside."""
    Z, Y, X = x.shape[:3]
    out = np.zeros_like(x)
    zs, ys, xs = slice(max(0, -dz), Z - max(0, dz)), slice(max(0, -dy), Y - max(0, dy)), \
        slice(max(0, -dx), X - max(0, dx))
    zt, yt, xt = slice(max(0, dz), Z - max(0, -dz)), slice(max(0, dy), Y - max(0, -dy)), \
        slice(max(0, dx), X - max(0, -dx))
    out[zs, ys, xs] = x[zt, yt, xt]
    return out


def conv3(x, w, b=None):
    """'same' 3x3x3 cross-correlation, x (Z,Y,X,Ci), w (27,Ci,Co), b (Co,) -> (Z,Y,X,Co)."""
    x = np.asarray(x, np.float64)
    y = np.zeros(x.shape[:3] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(OFFSETS):
        y += _shifted(x, dz, dy, dx) @ w[t]
    if b is not None:
        y += b
    return y


def conv3_backward(x, w, dy):
    """Gradients of conv3 for output gradient dy: (dx, dw, db)."""
    x = np.asarray(x, np.float64)
    dx = np.zeros_like(x)
    dw = np.zeros(w.shape)
    for t, (oz, oy, ox) in enumerate(OFFSETS):
        xs = _shifted(x, oz, oy, ox)
        dw[t] = xs.reshape(-1, x.shape[3]).T @ dy.reshape(-1, dy.shape[3])
        dx += _shifted(dy @ w[t].T, -oz, -oy, -ox)
    db = dy.reshape(-1, dy.shape[3]).sum(axis=0)
    return dx, dw, db


def init_params(seed=0, channels=CHANNELS, scale=1.0):
    """He-normal weights (last layer scaled down), zero biases."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(len(channels) - 1):
        ci, co = channels[i], channels[i + 1]
        std = np.sqrt(2.0 / (27 * ci)) * (0.1 if i == len(channels) - 2 else 1.0) * scale
        params.append((rng.standard_normal((27, ci, co)) * std, np.zeros(co)))
    return params


def forward(params, x, act_round=None):
    """pred = x + f(x) and the per-layer inputs / pre-activations for backward.

    act_round, if given, is applied to every stored hidden activation (e.g.
    ``bf16`` to follow a kernel that keeps activations in bfloat16)."""
    acts, pres = [np.asarray(x, np.float64)], []
    h = acts[0]
    for i, (w, b) in enumerate(params):
        z = conv3(h, w, b)
        pres.append(z)
        h = np.maximum(z, 0.0) if i < len(params) - 1 else z
        if i < len(params) - 1:
            if act_round is not None:
                h = act_round(h)
            acts.append(h)
    return acts[0] + h, (acts, pres)


def mse(pred, target):
    d = pred - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(params, cache, dpred, grad_round=None):
    """Parameter gradients for dL/dpred (the residual path adds nothing to them).

    grad_round, if given, is applied to each layer's output gradient before
    it is used (as a kernel that stores gradients in bfloat16 does)."""
    acts, pres = cache
    grads = [None] * len(params)
    rnd = grad_round if grad_round is not None else (lambda a: a)
    g = rnd(dpred)
    for i in range(len(params) - 1, -1, -1):
        w, _ = params[i]
        dx, dw, db = conv3_backward(acts[i], w, g)
        grads[i] = (dw, db)
        if i > 0:
            g = rnd(dx * (pres[i - 1] > 0))
    return grads


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
