"""FP64 CPU restatement of the Voigt-surrogate MLP forward and backward.

TEST INFRASTRUCTURE ONLY (checker of the HIP training kernels).

Mirrors the device arithmetic: the operands of every BF16-MFMA product (the
hidden weights, the hidden activations, and the pre-activation gradients
fed to the weight- and activation-gradient products) are rounded to BF16
with round-to-nearest-even; all accumulation is FP64 here (FP32 on the
device).  Loss = mean (out - t)^2 over the batch.
"""

from __future__ import annotations

import numpy as np

H = 64


def bf16(a):
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + 2 * H].reshape(H, 2); o += 2 * H
    b1 = p[o:o + H]; o += H
    Ws, bs = [], []
    for _ in range(3):
        Ws.append(p[o:o + H * H].reshape(H, H)); o += H * H
        bs.append(p[o:o + H]); o += H
    W5 = p[o:o + H]; o += H
    b5 = p[o]
    return W1, b1, Ws, bs, W5, b5


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def inputs(x, ly):
    u = 2.0 * np.arcsinh(np.abs(np.asarray(x, np.float32).astype(np.float64))) * float(np.float32(0.21714253)) - 1.0
    v = ((np.asarray(ly, np.float32) + np.float32(1.0)) * np.float32(1.0 / 3.0)).astype(np.float64)
    return u, v


def _ident(a):
    return np.asarray(a, np.float64)


def forward(p, x, ly, quantize=True):
    q = bf16 if quantize else _ident
    W1, b1, Ws, bs, W5, b5 = unpack(p)
    u, v = inputs(x, ly)
    z = [W1[:, 0:1] * u[None, :] + W1[:, 1:2] * v[None, :] + b1[:, None]]   # [64, n]
    h = [q(silu(z[0]))]
    for W, b in zip(Ws, bs):
        z.append(q(W) @ h[-1] + b[:, None])
        h.append(q(silu(z[-1])))
    out = W5 @ h[-1] + b5
    return out, z, h, u, v


def loss_and_grad(p, x, ly, t, quantize=True):
    """(sum of squared errors, gradient of the mean squared error).  With
    quantize=False every BF16 rounding is dropped (exact FP64 backprop)."""
    bf16_ = bf16 if quantize else _ident
    W1, b1, Ws, bs, W5, b5 = unpack(p)
    out, z, h, u, v = forward(p, x, ly, quantize)
    n = out.size
    err = out - t
    g = 2.0 * err / n
    grads = {}
    grads["b5"] = g.sum()
    grads["W5"] = h[3] @ g
    dz = W5[:, None] * g[None, :] * dsilu(z[3])
    gW, gb = [None] * 3, [None] * 3
    for layer in (2, 1, 0):
        gb[layer] = dz.sum(1)
        gW[layer] = bf16_(dz) @ h[layer].T
        dh = bf16_(Ws[layer]).T @ bf16_(dz)
        dz = dh * dsilu(z[layer])
    gW1 = np.stack([dz @ u, dz @ v], axis=1)
    gb1 = dz.sum(1)
    flat = np.concatenate([gW1.ravel(), gb1] + sum([[gW[i].ravel(), gb[i]] for i in range(3)], [])
                          + [grads["W5"], [grads["b5"]]])
    return float(np.sum(err * err)), flat
