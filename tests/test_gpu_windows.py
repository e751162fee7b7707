"""Benchmark workloads on the GPU against the oracle's golden window results
(tests/golden/windows.json, assembled by make_golden.py from per-part oracle
runs): every counter, the raw-set multiset digest and the maximal-set digest
must match."""
import hashlib
import json
import os
import sys

import pytest

import cuqc_amd as qc
from cuqc_amd import workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "windows.json")
WINDOWS = json.load(open(GOLDEN)) if os.path.exists(GOLDEN) else {}
KEYS = ["tasks", "branches", "children", "raw_cliques", "w_adj", "sum_t", "sum_t_written", "sum_clique_len", "beyond"]


def _digest(sets):
    h = hashlib.sha256()
    for s in sets:
        h.update((" ".join(str(int(x)) for x in s) + "\n").encode())
    return h.hexdigest()


def _check(res, gold):
    for k in KEYS:
        assert res.stats[k] == gold["stats"][k], k
    assert _digest(sorted(tuple(sorted(int(x) for x in s)) for s in res.raw)) == gold["raw_sha256"]
    assert len(res.maximal) == gold["maximal"]
    assert _digest(res.maximal_sorted_tuples()) == gold["maximal_sha256"]


@pytest.mark.parametrize("key", sorted(WINDOWS))
@pytest.mark.parametrize("order", ["fifo", "cost"])
def test_workload_window_persistent(key, order):
    gold = WINDOWS[key]
    w = workloads.WORKLOADS[int(key.split(":")[0])]
    g = workloads.graph(w, gold["instance"])
    assert hex(g.checksum()) == gold["checksum"]
    model = qc.CostModel.create(seed=1) if order == "cost" else None
    print(f"[gpu] workload {w.name} window {w.window} {order}", flush=True)
    sys.stdout.flush()
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", window=w.window, order=order, cost_model=model)
    _check(res, gold)
