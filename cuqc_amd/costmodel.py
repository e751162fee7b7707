"""Reference math of the task-cost MLP on the CPU (float64) and helpers to
turn task traces into training data.  The GPU kernels in csrc/mlp.hip are
checked against `forward` / `grad` here; `grad` itself is checked against
finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

NF, NH = 12, 64
OW1 = 0
OB1 = OW1 + NF * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NP = OB3 + 1


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(NF, NH), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3], p[OW3:OB3], p[OB3])


def standardize(feats, norm) -> np.ndarray:
    norm = np.asarray(norm, np.float64)
    return (np.log1p(np.maximum(np.asarray(feats, np.float64), 0.0)) - norm[:NF]) / norm[NF:2 * NF]


def forward(z: np.ndarray, p: np.ndarray):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    a1 = z @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3 + b3, (h1, h2)


def loss_and_grad(z: np.ndarray, t: np.ndarray, p: np.ndarray):
    """Mean squared error on standardized targets and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (h1, h2) = forward(z, p)
    n = len(t)
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = np.zeros(NP)
    g[OB3] = dy.sum()
    g[OW3:OB3] = h2.T @ dy
    d2 = np.outer(dy, W3) * (h2 > 0)
    g[OB2:OW3] = d2.sum(0)
    g[OW2:OB2] = (h1.T @ d2).ravel()
    d1 = (d2 @ W2.T) * (h1 > 0)
    g[OB1:OW2] = d1.sum(0)
    g[OW1:OB1] = (z.T @ d1).ravel()
    return loss, g


def subtree_sizes(parent: np.ndarray) -> np.ndarray:
    """Tasks per subtree from parent links (children after their parent)."""
    sub = np.ones(len(parent), np.int64)
    for i in range(len(parent) - 1, -1, -1):
        if parent[i] >= 0:
            sub[parent[i]] += sub[i]
    return sub
