"""SetList (raw sets as two arrays) and the order-independent multiset hash
the autotuner compares schedules with."""
from __future__ import annotations

import numpy as np
import pytest

from cuqc_amd import SetList, multiset_hash


def _sl(sets):
    off = np.zeros(len(sets) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in sets])
    v = np.array([x for s in sets for x in s], np.uint32)
    return SetList(off, v)


def test_sequence_behaviour():
    s = _sl([[1, 2, 3], [7, 4], [9]])
    assert len(s) == 3
    assert [list(map(int, x)) for x in s] == [[1, 2, 3], [7, 4], [9]]
    assert list(map(int, s[-1])) == [9]
    assert [list(map(int, x)) for x in s[1:]] == [[7, 4], [9]]
    with pytest.raises(IndexError):
        s[3]
    assert len(_sl([])) == 0


def test_hash_ignores_order_and_sees_every_change():
    rng = np.random.default_rng(5)
    sets = [sorted(rng.choice(1000, rng.integers(1, 12), replace=False).tolist()) for _ in range(300)]
    h = multiset_hash(_sl(sets))
    shuffled = [list(rng.permutation(s)) for s in sets]
    rng.shuffle(shuffled)
    assert multiset_hash(_sl(shuffled)) == h
    # one vertex changed, one set dropped, one set duplicated, two sets
    # trading a vertex (same vertex multiset overall): each changes the hash
    changed = [list(s) for s in sets]
    changed[17][0] = 1000 + changed[17][0]
    assert multiset_hash(_sl(changed)) != h
    assert multiset_hash(_sl(sets[1:])) != h
    assert multiset_hash(_sl(sets + [sets[3]])) != h
    a, b = [1, 2, 3], [4, 5]
    assert multiset_hash(_sl([a, b])) != multiset_hash(_sl([[1, 2, 4], [3, 5]]))
