"""Float64 restatement of the matching-order network declared in include/bms2.h (the
32-64-64-64-10 ReLU MLP, its loss, gradient, Adam update and bf16 inference).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the HIP kernels; the product never
imports it.  The network has no counterpart in the reference (its join fixes identity order,
isomorphism.hpp:371-447), so this follows the header's definitions; its gradient is pinned by
central finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, TARGETS = 32, 64, 10, 4
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (H, OUT)), ("b4", (OUT,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(s)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def round_bf16(x: np.ndarray) -> np.ndarray:
    """Round f32 values to bf16 (nearest, ties to even), returned as f32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    u = np.where(finite, u + 0x7FFF + ((u >> 16) & 1), u) & 0xFFFF0000
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def forward(p: np.ndarray, x: np.ndarray, bf16: bool = False):
    """(outputs [rows][10], hidden activations [h1, h2, h3])."""
    w = unpack(p)
    if bf16:
        w = {k: round_bf16(v).astype(np.float64) for k, v in w.items()}
        a = round_bf16(x).astype(np.float64)
    else:
        a = np.asarray(x, np.float64)
    hs = []
    for i in (1, 2, 3):
        a = np.maximum(a @ w[f"W{i}"] + w[f"b{i}"], 0.0)
        if bf16:
            a = round_bf16(a).astype(np.float64)
        hs.append(a)
    return a @ w["W4"] + w["b4"], hs


def min_abs_preactivation(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Per row, the smallest |pre-activation| over the hidden units: rows within f32 rounding
    of a ReLU kink have a gradient that f32 and f64 may legitimately disagree on."""
    w = unpack(p)
    a = np.asarray(x, np.float64)
    m = np.full(a.shape[0], np.inf)
    for i in (1, 2, 3):
        z = a @ w[f"W{i}"] + w[f"b{i}"]
        m = np.minimum(m, np.abs(z).min(axis=1))
        a = np.maximum(z, 0.0)
    return m


def loss_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, depth_weight: float = 1.0):
    """Loss of bms2.h and its gradient with respect to the flat parameters."""
    w = unpack(p)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    y, (h1, h2, h3) = forward(p, x)
    wr = t[:, 3] / t[:, 3].sum()
    e0, e1 = y[:, 0] - t[:, 0], y[:, 1] - t[:, 1]
    tk = np.clip(t[:, 2].astype(np.int64), 0, 7)
    z = y[:, 2:]
    zm = z.max(axis=1, keepdims=True)
    ez = np.exp(z - zm)
    sm = ez / ez.sum(axis=1, keepdims=True)
    ce = np.log(ez.sum(axis=1)) + zm[:, 0] - z[np.arange(len(z)), tk]
    loss = float(np.sum(wr * (e0 ** 2 + e1 ** 2 + depth_weight * ce)))
    dy = np.zeros_like(y)
    dy[:, 0] = 2 * e0 * wr
    dy[:, 1] = 2 * e1 * wr
    onehot = np.zeros_like(sm)
    onehot[np.arange(len(z)), tk] = 1.0
    dy[:, 2:] = depth_weight * wr[:, None] * (sm - onehot)
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    dz3 = (dy @ w["W4"].T) * (h3 > 0)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ w["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ dz1, dz1.sum(0)
    return loss, pack(g)


class Adam:
    """Adam with bias correction: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)."""

    def __init__(self, n: int, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps

    def step(self, p: np.ndarray, g: np.ndarray) -> np.ndarray:
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)
