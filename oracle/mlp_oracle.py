"""CPU oracle of the closure MLP (test infrastructure only).

Restates, in numpy, the network of exactdiag_amd/csrc/ed_mlp.hip: dense
layers with GELU (tanh form) on hidden layers, mean-squared-error loss, BF16
rounding (round to nearest even) of every matrix-product operand, FP32
elsewhere.  ``exact=True`` runs the same formulas in FP64 without rounding;
tests pin that variant by central finite differences, and the BF16 variant
is the reference for the GPU kernels' outputs and gradients.
"""

from __future__ import annotations

import numpy as np

K_GELU = 0.7978845608028654


def bf16(a):
    """Round float32 values to bfloat16 (nearest even) and return them as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    return np.where(nan, a, out)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x**3)))


def gelu_grad(x):
    t = np.tanh(K_GELU * (x + 0.044715 * x**3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def unpack(params, dims):
    out, off = [], 0
    for l in range(1, len(dims)):
        k, n = dims[l - 1], dims[l]
        W = params[off: off + k * n].reshape(k, n)
        off += k * n
        b = params[off: off + n]
        off += n
        out.append((W, b))
    return out


def forward(params, dims, x, exact=False):
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dt)) if exact else bf16
    layers = unpack(np.asarray(params, dt), dims)
    a16 = rnd(x)
    zs, ins = [], []
    for l, (W, b) in enumerate(layers):
        ins.append(a16)
        z = (a16.astype(np.float64) @ rnd(W).astype(np.float64)).astype(dt) + b
        zs.append(z)
        if l < len(layers) - 1:
            a16 = rnd(gelu(z).astype(dt))
    return zs, ins


def loss_and_grad(params, dims, x, t, exact=False):
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dt)) if exact else bf16
    layers = unpack(np.asarray(params, dt), dims)
    zs, ins = forward(params, dims, x, exact)
    y = zs[-1]
    n = y.size
    d = (y - t).astype(dt)
    loss = float(np.sum(d.astype(np.float64) ** 2) / n)
    dz = (2.0 * d / n).astype(dt)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dz16 = rnd(dz)
        dW = (ins[l].astype(np.float64).T @ dz16.astype(np.float64)).astype(dt)
        db = np.sum(dz, axis=0, dtype=np.float64).astype(dt)
        grads.append((dW, db))
        if l > 0:
            da = (dz16.astype(np.float64) @ rnd(W).astype(np.float64).T).astype(dt)
            dz = (da * gelu_grad(zs[l - 1])).astype(dt)
    grads.reverse()
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat


def init_params(dims, seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(1, len(dims)):
        parts.append(rng.normal(0, np.sqrt(2.0 / dims[l - 1]), size=dims[l - 1] * dims[l]))
        parts.append(rng.normal(0, 0.1, size=dims[l]))
    return np.concatenate(parts).astype(np.float32)
