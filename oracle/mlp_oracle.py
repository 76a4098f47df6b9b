"""CPU oracle of the learned-light network (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker of the GPU
network kernels; the product package never imports it.

The network has no counterpart in the reference: it predicts the optical
depth of GetLightEnergy's cone march (cloudRayMarch.comp:338-373).  Layout
of the FP32 parameter vector (shared with include/meteoros_cloud.h):
W1[9][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4
with W[k][n] the weight from input k to output n.

Two modes:
  quantized=True   the GPU's arithmetic: inputs, W1..W3 and the hidden
                   activations rounded to BF16 (round to nearest even);
                   biases, w4, b4 and all sums in FP64 (the GPU sums in FP32)
  quantized=False  the plain FP64 network, used to pin the backward pass
                   against central finite differences
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 9, 64
OFF = {"W1": (0, (9, 64)), "b1": (576, (64,)), "W2": (640, (64, 64)), "b2": (4736, (64,)),
       "W3": (4800, (64, 64)), "b3": (8896, (64,)), "w4": (8960, (64,)), "b4": (9024, ())}
N_PARAMS = 9025


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    assert p.size == N_PARAMS
    out = {}
    for k, (o, shape) in OFF.items():
        n = int(np.prod(shape)) if shape else 1
        out[k] = p[o:o + n].reshape(shape) if shape else p[o]
    return out


def pack(d: dict) -> np.ndarray:
    p = np.zeros(N_PARAMS, dtype=np.float64)
    for k, (o, shape) in OFF.items():
        v = np.asarray(d[k], dtype=np.float64).reshape(-1)
        p[o:o + v.size] = v
    return p


def bf16(a) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def silu_grad(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(p: np.ndarray, x: np.ndarray, quantized: bool = True, keep: bool = False):
    """Network output for inputs x[n, 9]; with keep=True also the
    intermediate (a0, z1, h1, z2, h2, z3, h3) for backward()."""
    d = unpack(p)
    q = bf16 if quantized else (lambda v: np.asarray(v, dtype=np.float64))
    a0 = q(x)
    z1 = a0 @ q(d["W1"]) + d["b1"]
    h1 = q(silu(z1))
    z2 = h1 @ q(d["W2"]) + d["b2"]
    h2 = q(silu(z2))
    z3 = h2 @ q(d["W3"]) + d["b3"]
    h3 = q(silu(z3))
    y = h3 @ d["w4"] + d["b4"]
    if keep:
        return y, (a0, z1, h1, z2, h2, z3, h3)
    return y


def loss_and_grad(p: np.ndarray, x: np.ndarray, y: np.ndarray, quantized: bool = True):
    """Mean squared error and its gradient w.r.t. the parameter vector.
    Rounding is treated as the identity in the backward pass, as the GPU
    training step does."""
    d = unpack(p)
    q = bf16 if quantized else (lambda v: np.asarray(v, dtype=np.float64))
    yhat, (a0, z1, h1, z2, h2, z3, h3) = forward(p, x, quantized, keep=True)
    n = x.shape[0]
    err = yhat - np.asarray(y, dtype=np.float64)
    loss = float(np.mean(err * err))
    dy = 2.0 * err / n
    g = {"w4": h3.T @ dy, "b4": dy.sum()}
    dz3 = np.outer(dy, d["w4"]) * silu_grad(z3)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ q(d["W3"]).T) * silu_grad(z2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ q(d["W2"]).T) * silu_grad(z1)
    g["W1"], g["b1"] = a0.T @ dz1, dz1.sum(0)
    return loss, pack(g)


def mse(p: np.ndarray, x: np.ndarray, y: np.ndarray, quantized: bool = False) -> float:
    e = forward(p, x, quantized) - np.asarray(y, dtype=np.float64)
    return float(np.mean(e * e))
