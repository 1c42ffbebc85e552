"""CPU oracle of the learned subgrid CNN (no reference counterpart; the
specification is this file).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Network: layers l = 0..L-1 of periodic 3x3x3 convolutions
    z_l[n, co] = b_l[co] + sum_{tap, ci} a_l[n + off(tap), ci] W_l[tap, co, ci]
with off(tap) = (tap // 9 - 1, tap // 3 % 3 - 1, tap % 3 - 1) on an (N, N, N)
periodic grid, a_{l+1} = relu(z_l) for l < L-1, output y = z_{L-1}.
Loss: mean over nodes and output channels of (y - t)^2.

``bf16=True`` reproduces the rounding points of the GPU kernels: inputs,
weights, hidden activations and back-propagated gradients are rounded to
bfloat16 (round to nearest even); sums are exact (float64), the GPU's are
FP32.  ``bf16=False`` is plain float64 arithmetic (the finite-difference
check uses it).
"""
from __future__ import annotations

import numpy as np

OFFSETS = [(t // 9 - 1, t // 3 % 3 - 1, t % 3 - 1) for t in range(27)]


def to_bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def shift(a: np.ndarray, off) -> np.ndarray:
    """b[n] = a[n + off] on the periodic grid (axes 0..2)."""
    return np.roll(a, shift=(-off[0], -off[1], -off[2]), axis=(0, 1, 2))


def conv(a: np.ndarray, W: np.ndarray, b=None) -> np.ndarray:
    """a (N, N, N, ci), W (27, co, ci) -> (N, N, N, co)."""
    out = np.zeros(a.shape[:3] + (W.shape[1],))
    for t, off in enumerate(OFFSETS):
        out += shift(a, off) @ W[t].T
    if b is not None:
        out += b
    return out


def forward(layers, x, bf16=True):
    """layers: [(W (27, co, ci), b (co,))]; x (N, N, N, c0).  Returns (y, acts)
    with acts[l] the input of layer l."""
    rd = to_bf16 if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    a = rd(x)
    acts = [a]
    y = None
    for l, (W, b) in enumerate(layers):
        z = conv(a, rd(W), np.asarray(b, dtype=np.float64))
        if l + 1 < len(layers):
            a = rd(np.maximum(z, 0.0))
            acts.append(a)
        else:
            y = z
    return y, acts


def loss_and_grads(layers, x, t, bf16=True):
    """Mean-squared-error loss and its gradients [(dW, db)] per layer."""
    rd = to_bf16 if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y, acts = forward(layers, x, bf16)
    n = y.size
    d = y - t
    loss = float(np.sum(d * d) / n)
    dz = rd(2.0 * d / n)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        a = acts[l]
        dW = np.zeros(W.shape)
        for tp, off in enumerate(OFFSETS):
            sa = shift(a, off)
            dW[tp] = np.einsum("xyzo,xyzi->oi", dz, sa)
        grads[l] = (dW, dz.sum(axis=(0, 1, 2)))
        if l > 0:
            Wr = rd(W)
            da = np.zeros(a.shape)
            for tp, off in enumerate(OFFSETS):
                da += shift(dz, (-off[0], -off[1], -off[2])) @ Wr[tp]
            dz = rd(da * (acts[l] > 0))
    return loss, grads


def box_filter(u: np.ndarray) -> np.ndarray:
    """Coarse field of the deconvolution task: at coarse node i (fine node 2i),
    the separable (1/4, 1/2, 1/4) average of the fine field (N, N, N, c)."""
    w = (0.25, 0.5, 0.25)
    out = np.zeros_like(u)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                out += w[a + 1] * w[b + 1] * w[c + 1] * shift(u, (a, b, c))
    return out[::2, ::2, ::2]


def forward_at(layers, x, nodes, bf16=True):
    """The network output at the listed nodes only (nodes (S, 3) int), with the
    rounding points of ``forward``: each output depends on the (2L+1)^3 input
    neighbourhood of its node, which is cut out periodically and convolved
    without padding, layer by layer.  Equals ``forward(layers, x)[0][nodes]``."""
    rd = to_bf16 if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    L = len(layers)
    N = x.shape[0]
    r = np.arange(-L, L + 1)
    out = []
    for n in np.asarray(nodes):
        ix = [(n[d] + r) % N for d in range(3)]
        a = rd(x[np.ix_(ix[0], ix[1], ix[2])])
        for l, (W, b) in enumerate(layers):
            m = a.shape[0] - 2
            Wr = rd(W)
            z = np.zeros((m, m, m, W.shape[1])) + np.asarray(b, dtype=np.float64)
            for t, off in enumerate(OFFSETS):
                z += a[1 + off[0]: 1 + off[0] + m, 1 + off[1]: 1 + off[1] + m, 1 + off[2]: 1 + off[2] + m] @ Wr[t].T
            a = rd(np.maximum(z, 0.0)) if l + 1 < L else z
        out.append(a[0, 0, 0])
    return np.array(out)
