"""CPU reference of the warm-start network (test infrastructure only).

The network has no counterpart in the reference; this restates what the
device kernels compute, in float64 numpy with the device's BF16 roundings
applied at the same points, so the GPU's forward pass, loss and gradients
can be checked. BF16 operands: the standardized inputs, the weights, the
hidden activations, and the backward pass's output and hidden gradients.
Accumulation is exact here and FP32 on the device, so checks use norms with
tolerances, not bit equality.
"""
from __future__ import annotations

import numpy as np


def bf16(a) -> np.ndarray:
    """FP32 -> BF16 -> FP32, round to nearest even (as the device's f2bf)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params, dims):
    """Layer by layer: W [out][in] row-major, then b [out]."""
    out, off = [], 0
    for i, o in zip(dims[:-1], dims[1:]):
        W = params[off:off + i * o].reshape(o, i).astype(np.float64)
        off += i * o
        b = params[off:off + o].astype(np.float64)
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def forward(params, dims, inputs, in_mean, in_scale):
    layers = unpack(params, dims)
    x = bf16(((inputs - in_mean) / in_scale).astype(np.float32))
    xs, zs = [x], []
    for li, (W, b) in enumerate(layers):
        z = x @ bf16(W).T + b
        zs.append(z)
        if li + 1 < len(layers):
            x = bf16(silu(z))
            xs.append(x)
    return xs, zs


def loss_and_grads(params, dims, inputs, targets, in_mean, in_scale, out_mean, out_scale):
    """targets [rows][n_out + 1] with ground first (as esl_op_solve writes)."""
    layers = unpack(params, dims)
    xs, zs = forward(params, dims, inputs, in_mean, in_scale)
    t = ((targets[:, 1:] - out_mean) / out_scale).astype(np.float32).astype(np.float64)
    y = zs[-1]
    rows, n_out = y.shape
    d = y - t
    loss = (d * d).sum() / (rows * n_out)
    dz = 2.0 * d / (rows * n_out)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        dzb = bf16(dz.astype(np.float32))
        gW = dzb.T @ xs[li]
        gb = dz.sum(0)
        grads.append((gW, gb))
        if li == 0:
            break
        dh = dzb @ bf16(W)
        dz = dh * silu_grad(zs[li - 1])
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def predict(params, dims, inputs, in_mean, in_scale, out_mean, out_scale):
    _, zs = forward(params, dims, inputs, in_mean, in_scale)
    y = zs[-1]
    return np.concatenate([np.zeros((y.shape[0], 1)), y * out_scale + out_mean], axis=1)
