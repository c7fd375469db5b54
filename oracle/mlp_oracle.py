"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

CPU restatement of the learned-Q network (lulesh-accelerate_amd/csrc/
lulesh_mlp.hpp, lulesh_mlp.hip): a 12-64-64-2 perceptron with ReLU hidden
layers.  Two forms:

  forward64 / grad64   float64 arithmetic without any rounding: the
                       mathematical network, pinned by central finite
                       differences in the tests;
  forward_bf16 /       the device arithmetic: standardised inputs, weights
  grad_bf16            and activations rounded to bf16 (round to nearest
                       even) where the device rounds them, products
                       accumulated in float32.

Parameter vector layout: W1[12][64] (input-major), b1[64], W2[64][64],
b2[64], W3[64][2], b3[2].  Loss: 0.5 / N * sum_s sum_o w_o (y_so - t_so)^2.
Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 12, 64, 2
SIZES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)),
         ("b2", (N_HID,)), ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n]).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    """He-scaled normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / N_IN), (N_IN, N_HID)),
         "b1": np.zeros(N_HID), "W2": rng.normal(0, np.sqrt(2 / N_HID), (N_HID, N_HID)),
         "b2": np.zeros(N_HID), "W3": rng.normal(0, np.sqrt(1 / N_HID), (N_HID, N_OUT)),
         "b3": np.zeros(N_OUT)}
    return pack(d).astype(np.float32)


def forward64(p, z):
    w = unpack(np.asarray(p, np.float64))
    a1 = z @ w["W1"] + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(a2, 0)
    return h2 @ w["W3"] + w["b3"], (z, a1, h1, a2, h2)


def loss64(p, z, t, wo=(1.0, 1.0)):
    y, _ = forward64(p, z)
    return 0.5 * np.sum(np.asarray(wo) * (y - t) ** 2) / len(z)


def grad64(p, z, t, wo=(1.0, 1.0)):
    w = unpack(np.asarray(p, np.float64))
    y, (z, a1, h1, a2, h2) = forward64(p, z)
    dy = np.asarray(wo) * (y - t) / len(z)
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"].T) * (a2 > 0)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ w["W2"].T) * (a1 > 0)
    g["W1"], g["b1"] = z.T @ d1, d1.sum(0)
    return pack(g)


def standardise(x, mean, std):
    return ((np.asarray(x, np.float32) - np.float32(mean)) *
            (np.float32(1) / np.asarray(std, np.float32))).astype(np.float32)


def forward_bf16(p, x, mean, std):
    """Device-rounded forward pass; returns outputs and the bf16 tensors."""
    w = unpack(np.asarray(p, np.float32))
    z = bf16(standardise(x, mean, std))
    a1 = z @ bf16(w["W1"]) + w["b1"]
    h1 = bf16(np.maximum(a1, 0))
    a2 = h1 @ bf16(w["W2"]) + w["b2"]
    h2 = bf16(np.maximum(a2, 0))
    y = h2 @ bf16(w["W3"]) + w["b3"]
    return y.astype(np.float32), (z, h1, h2)


def grad_bf16(p, x, t, mean, std, wo=(1.0, 1.0)):
    """Device-rounded gradient: the operands of every product are rounded
    to bf16 as on the matrix cores (dH2 is formed in fp32 from the fp32
    W3, as on the device)."""
    w = unpack(np.asarray(p, np.float32))
    y, (z, h1, h2) = forward_bf16(p, x, mean, std)
    n = len(x)
    dy = (np.asarray(wo, np.float32) * (y - t) / np.float32(n)).astype(np.float32)
    g = {"W3": h2.T @ bf16(dy), "b3": dy.sum(0)}
    d2 = ((dy @ w["W3"].T) * (h2 > 0)).astype(np.float32)
    g["W2"], g["b2"] = h1.T @ bf16(d2), d2.sum(0)
    d1 = ((bf16(d2) @ bf16(w["W2"]).T) * (h1 > 0)).astype(np.float32)
    g["W1"], g["b1"] = z.T @ bf16(d1), d1.sum(0)
    loss = 0.5 * np.sum(np.asarray(wo) * (y.astype(np.float64) - t) ** 2) / n
    return pack(g).astype(np.float32), loss
