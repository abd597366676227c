"""CPU reference of the surrogate network (numpy, float64).

TEST INFRASTRUCTURE ONLY: used by tests/ (and the surrogate's evaluation
scripts under tests/) to check the HIP MLP kernels of axonexa_amd/csrc/mlp.hip;
the product never imports it.

Network: hidden layers z = W h + b, h' = GELU(z) (tanh form), last layer
linear; loss = mean over rows and outputs of (y - t)^2.  Parameters are
packed like the C ABI (per layer W[out][in] row major, then b[out]).

With ``bf16=True`` the forward pass rounds what the GPU keeps in bf16 (the
inputs, the weights, the pre-activations and activations between layers) to
bf16 (round to nearest even), so the remaining GPU/CPU differences are fp32
accumulation order and the GPU's bf16 gradients.
"""
from __future__ import annotations

import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)


def bf16_round(a) -> np.ndarray:
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * x * x)


def unpack(params: np.ndarray, dims):
    out, o = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = params[o:o + a * b].reshape(b, a)
        o += a * b
        bias = params[o:o + b]
        o += b
        out.append((np.asarray(W, np.float64), np.asarray(bias, np.float64)))
    assert o == params.size
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def init_params(dims, seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(in), 1/sqrt(in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    layers = []
    for a, b in zip(dims[:-1], dims[1:]):
        lim = 1.0 / np.sqrt(a)
        layers.append((rng.uniform(-lim, lim, size=(b, a)), np.zeros(b)))
    return pack(layers).astype(np.float32)


def forward(params, dims, x, bf16: bool = False):
    rnd = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    layers = unpack(np.asarray(params, np.float64), dims)
    h = rnd(x)
    cache = []
    for k, (W, b) in enumerate(layers):
        Wr = rnd(W)
        z = h @ Wr.T + b
        if k + 1 < len(layers):
            zr = rnd(z)
            cache.append((h, Wr, zr))
            h = rnd(gelu(z))
        else:
            cache.append((h, Wr, None))
            return z, cache
    raise AssertionError("unreachable")


def loss_grad(params, dims, x, t, bf16: bool = False):
    """Mean squared error and its gradient (packed like params)."""
    y, cache = forward(params, dims, x, bf16)
    n, m = t.shape
    d = y - t
    loss = float(np.mean(d * d))
    g = 2.0 * d / (n * m)
    grads = [None] * len(cache)
    for k in range(len(cache) - 1, -1, -1):
        h, Wr, _ = cache[k]
        grads[k] = (g.T @ h, g.sum(axis=0))
        if k > 0:
            zprev = cache[k - 1][2]
            g = (g @ Wr) * gelu_grad(zprev)
    return loss, pack(grads)
