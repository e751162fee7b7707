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


def subtree_sizes(parent: np.ndarray, beyond: np.ndarray | None = None) -> np.ndarray:
    """Task records per subtree from parent links (children after their
    parent); ``beyond`` adds each task's children past a depth window, which
    are created but not expanded."""
    sub = np.ones(len(parent), np.int64)
    if beyond is not None:
        sub += np.asarray(beyond, np.int64)
    for i in range(len(parent) - 1, -1, -1):
        if parent[i] >= 0:
            sub[parent[i]] += sub[i]
    return sub


def trace_rows(results) -> tuple[np.ndarray, np.ndarray]:
    """Features and log subtree sizes of every traced task in ``results``
    (results of ``mine(..., trace=True)`` / ``mine_batch(..., trace=True)``)."""
    feats, target = [], []
    for r in results:
        if r.trace_feats is None or len(r.trace_feats) == 0:
            continue
        feats.append(np.asarray(r.trace_feats, np.float32))
        target.append(np.log(np.asarray(r.trace_subtree, np.float64)).astype(np.float32))
    if not feats:
        return np.zeros((0, NF), np.float32), np.zeros(0, np.float32)
    return np.concatenate(feats), np.concatenate(target)


def train_from_traces(jobs, steps: int = 400, lr: float = 3e-3, seed: int = 0, max_rows: int = 200_000):
    """Log traces of the given mining jobs with the FIFO persistent kernel and
    train the cost model on them with the GPU kernels.

    jobs: iterable of (graph, gamma, min_size, window) tuples.
    Returns (model, rows used)."""
    from . import CostModel, mine

    res = [mine(g, ga, ms, 0, mode="persistent", window=w, trace=True) for g, ga, ms, w in jobs]
    X, y = trace_rows(res)
    if len(y) == 0:
        raise ValueError("the traced jobs expanded no tasks")
    if len(y) > max_rows:
        idx = np.random.default_rng(seed).choice(len(y), max_rows, replace=False)
        X, y = X[idx], y[idx]
    return CostModel.train(X, y, steps=steps, lr=lr, seed=seed), len(y)
