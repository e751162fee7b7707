"""Small seeded graphs shared by the tests."""
from __future__ import annotations

import numpy as np


def csr_from_edges(n: int, edges) -> tuple[np.ndarray, np.ndarray]:
    e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 2)
    if len(e):
        e = e[e[:, 0] != e[:, 1]]
        e = np.unique(np.sort(e, axis=1), axis=0)
        both = np.concatenate([e, e[:, ::-1]])
        order = np.lexsort((both[:, 1], both[:, 0]))
        both = both[order]
    else:
        both = np.zeros((0, 2), np.int64)
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, both[:, 0] + 1, 1)
    return np.cumsum(off).astype(np.uint64), both[:, 1].astype(np.uint32)


def planted(n: int, p: float, sizes, keep: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """G(n, p) with dense planted groups (each internal pair kept w.p. `keep`)."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    mask = rng.random(len(iu[0])) < p
    edges = set(zip(iu[0][mask].tolist(), iu[1][mask].tolist()))
    perm = rng.permutation(n)
    at = 0
    for k in sizes:
        mem = sorted(perm[at:at + k].tolist())
        at += k
        for i in range(k):
            for j in range(i + 1, k):
                e = (mem[i], mem[j])
                if rng.random() < keep:
                    edges.add(e)
                else:
                    edges.discard(e)
    return csr_from_edges(n, edges)


def complete(k: int):
    return csr_from_edges(k, [(i, j) for i in range(k) for j in range(i + 1, k)])


# (name, builder, gamma, min_size)
SMALL_CASES = [
    ("er60_p30_g07", lambda: planted(60, 0.30, [], 1.0, 11), 0.7, 6),
    ("er80_planted_g09", lambda: planted(80, 0.10, [12, 15], 0.95, 12), 0.9, 8),
    ("er120_planted_g085", lambda: planted(120, 0.08, [14, 18, 22], 0.92, 13), 0.85, 10),
    ("er40_dense_g06", lambda: planted(40, 0.45, [], 1.0, 14), 0.6, 7),
    ("er150_planted_g075", lambda: planted(150, 0.06, [10, 16, 20, 25], 0.9, 15), 0.75, 9),
    ("er200_planted_g095", lambda: planted(200, 0.05, [30, 40], 0.98, 16), 0.95, 12),
    ("er90_g05", lambda: planted(90, 0.2, [10], 0.9, 17), 0.5, 10),
    ("k12", lambda: complete(12), 0.9, 5),
]
