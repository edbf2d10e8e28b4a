"""CPU reference of the cooling-surrogate MLP (TEST INFRASTRUCTURE ONLY).

The network replaces coolingh's per-cell update (src/coolingh.f90:171-259,
`atomic`): 4 -> 64 -> 64 -> 64 -> 2 with SiLU after each hidden layer.
This module restates the forward pass, the backward pass of the mean
squared error and one Adam update in numpy, in FP64, optionally rounding to
bfloat16 at exactly the points where the gfx950 kernels hold bf16 values
(inputs, weights and hidden activations that feed the matrix cores), so that
the kernels' outputs and gradients can be checked against it.  The backward
pass is pinned independently by central finite differences of the FP64
forward pass (tests/test_surrogate_ref.py).

Parameter vector layout (float32 on the device, 8770 values):
    W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[2][64], b4[2]
(row-major, out x in), in that order.
"""
from __future__ import annotations

import numpy as np

SIZES = [(64, 4), (64, 64), (64, 64), (2, 64)]
NPARAM = sum(o * i + o for o, i in SIZES)   # 8770


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Parameter vector -> [(W, b), ...] float64 views."""
    theta = np.asarray(theta, dtype=np.float64)
    out, k = [], 0
    for o, i in SIZES:
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def init_params(seed=0):
    """Kaiming-uniform-like initialisation (float32 values)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in SIZES:
        lim = np.sqrt(6.0 / i) * 0.5
        parts.append(rng.uniform(-lim, lim, size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x, round_bf16=True):
    """x: [S][4] standardised inputs -> (y [S][2], cache)."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(theta)
    a = r(x)
    acts, zs = [a], []
    for li, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        zs.append(z)
        if li < 3:
            a = r(silu(z))
            acts.append(a)
        else:
            a = z
    return a, (acts, zs)


def loss_and_grad(theta, x, t, w=(1.0, 1.0), round_bf16=True):
    """Mean over samples of sum_j w_j (y_j - t_j)^2 and its gradient with
    respect to the parameter vector (the quantity the training kernel
    accumulates)."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    S = x.shape[0]
    y, (acts, zs) = forward(theta, x, round_bf16)
    w = np.asarray(w, dtype=np.float64)
    L = float(np.mean(np.sum(w * (y - t) ** 2, axis=1)))
    layers = unpack(theta)
    grads = [None] * 4
    d = 2.0 * w * (y - t) / S            # dL/dz4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        dd = r(d)                         # the kernel feeds bf16 deltas to the matrix cores
        gW = dd.T @ acts[li]
        gb = d.sum(axis=0)
        grads[li] = (gW, gb)
        if li > 0:
            da = dd @ r(W)
            d = da * dsilu(zs[li - 1])
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, g


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), float64 arithmetic on float32 state."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
