"""FP64 CPU oracle of the surrogate network (test infrastructure only: tests, smoke and the
bench's CPU leg load it, as the checker of libwovengpu.so's wt_mlp_* kernels).

The network has no counterpart in the reference. It is a multilayer perceptron
sizes[0] -> ... -> sizes[-1], GELU (exact, erf form) after each hidden layer, linear output,
masked mean squared error L = sum(m (y - t)^2) / sum(m), Adam (0.9, 0.999, 1e-8) with the
step lr * sqrt(1 - b2^t) / (1 - b1^t) * m / (sqrt(v) + eps).

`emulate=True` reproduces where the GPU rounds to BF16 (round to nearest even): the network
input, the weights used by the products, each hidden activation after GELU, and the
backpropagated dZ used by the products; the bias gradient uses the unrounded dZ. Everything
else is FP64 here (FP32 on the GPU).

Pins (tests/test_mlp_oracle.py): gradients against torch autograd in float64 and against
central finite differences; BF16 rounding against torch's float32 -> bfloat16 conversion; GELU
against erf closed forms; the Adam step against torch.optim.Adam with eps = 0.
"""
from __future__ import annotations

import math

import numpy as np

_SQRT1_2 = 1.0 / math.sqrt(2.0)
_INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)
_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x) -> np.ndarray:
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * _SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z * _SQRT1_2)) + z * _INV_SQRT_2PI * np.exp(-0.5 * z * z)


def unflatten(sizes, flat):
    """Flat order: per layer W[out][in] row-major, then b[out]."""
    flat = np.asarray(flat, dtype=np.float64)
    Ws, bs, off = [], [], 0
    for i, o in zip(sizes[:-1], sizes[1:]):
        Ws.append(flat[off:off + o * i].reshape(o, i))
        off += o * i
        bs.append(flat[off:off + o].copy())
        off += o
    if off != flat.size:
        raise ValueError("parameter count does not match the layer sizes")
    return Ws, bs


def flatten(Ws, bs) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def forward(sizes, flat, X, emulate: bool = True):
    """Returns (y, cache) with cache = (inputs of each layer, pre-activations)."""
    Ws, bs = unflatten(sizes, flat)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    a = r(X)
    ins, zs = [], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        ins.append(a)
        z = a @ r(W).T + b
        zs.append(z)
        if l < len(Ws) - 1:
            a = r(gelu(z))
    return zs[-1], (ins, zs)


def loss_grad(sizes, flat, X, T, mask=None, emulate: bool = True):
    """(loss, flat gradient)."""
    Ws, _ = unflatten(sizes, flat)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    y, (ins, zs) = forward(sizes, flat, X, emulate)
    T = np.asarray(T, dtype=np.float64)
    M = np.ones_like(T) if mask is None else np.asarray(mask, dtype=np.float64)
    msum = M.sum()
    loss = float((M * (y - T) ** 2).sum() / msum) if msum > 0 else 0.0
    dz = 2.0 * M * (y - T) / msum if msum > 0 else np.zeros_like(y)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        gW[l] = r(dz).T @ ins[l]
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        da = r(dz) @ r(Ws[l])
        dz = da * gelu_grad(zs[l - 1])
    return loss, flatten(gW, gb)


class Adam:
    """The library's Adam update on a flat FP64 parameter vector."""

    def __init__(self, n: int, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.b1, self.b2, self.eps = b1, b2, eps

    def step(self, w, g, lr: float):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        lr_t = lr * math.sqrt(1 - self.b2 ** self.t) / (1 - self.b1 ** self.t)
        return w - lr_t * self.m / (np.sqrt(self.v) + self.eps)


def init_params(sizes, seed: int = 0) -> np.ndarray:
    """Random parameters for tests (He-uniform weights, small random biases)."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for i, o in zip(sizes[:-1], sizes[1:]):
        a = math.sqrt(6.0 / i)
        Ws.append(rng.uniform(-a, a, (o, i)))
        bs.append(rng.uniform(-0.1, 0.1, o))
    return flatten(Ws, bs)
