"""NumPy FP64 restatement of the surrogate MLP (test infrastructure): forward, loss and
backward of n_in -> hidden... -> n_out with GELU (tanh form), sigmoid output and mean squared
error, the operands optionally rounded to BF16 the way the GPU kernels round them."""
from __future__ import annotations

import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)


def to_bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * z * (1 + t)


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * z * z)


def forward(layers, x, bf16=False):
    """layers: [(W [out][in], b)], x [rows][in].  Returns (yhat, cache)."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(x)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = h @ r(W).T + np.asarray(b, np.float64)
        cache.append((h, z))
        if li < len(layers) - 1:
            h = r(gelu(z).astype(np.float32)) if bf16 else gelu(z)
        else:
            h = 1.0 / (1.0 + np.exp(-z))
    return h, cache


def loss_and_grads(layers, x, y, bf16=False):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    yhat, cache = forward(layers, x, bf16)
    n = yhat.size
    loss = float(np.mean((yhat - y) ** 2))
    dz = 2.0 * (yhat - y) * yhat * (1 - yhat) / n
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        h, z = cache[li]
        dzr = r(dz.astype(np.float32)) if bf16 else dz
        grads[li] = (dzr.T @ h, dz.sum(axis=0))
        if li > 0:
            dh = dzr @ r(layers[li][0])
            dz = dh * gelu_grad(cache[li - 1][1])
    return loss, grads
