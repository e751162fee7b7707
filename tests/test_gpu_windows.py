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


SINGLE = sorted(k for k in WINDOWS if ":" not in k)
BATCHES = sorted({int(k.split(":")[0]) for k in WINDOWS if ":" in k})


@pytest.mark.parametrize("key", SINGLE)
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


@pytest.mark.parametrize("key", [k for k in ("2", "4") if k in WINDOWS])
def test_workload_window_plain(key):
    """The plain form on windows that start deep in the root's branches (the
    leading removals are applied to the root before the first level)."""
    gold = WINDOWS[key]
    w = workloads.WORKLOADS[int(key)]
    g = workloads.graph(w, gold["instance"])
    assert hex(g.checksum()) == gold["checksum"]
    print(f"[gpu] workload {w.name} window {w.window} plain", flush=True)
    sys.stdout.flush()
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="plain", window=w.window)
    _check(res, gold)


def _batch(wl):
    w = workloads.WORKLOADS[wl]
    golds = [WINDOWS[f"{wl}:{i}"] for i in range(w.instances)]
    gs = workloads.graphs(w)
    for g, gold in zip(gs, golds):
        assert hex(g.checksum()) == gold["checksum"]
    return w, gs, golds


@pytest.mark.parametrize("wl", BATCHES)
@pytest.mark.parametrize("order", ["fifo", "cost"])
def test_batch_workload_one_launch(wl, order):
    """Every instance of the batch in one persistent launch (hipGraph, two
    replays) equals its own oracle result."""
    w, gs, golds = _batch(wl)
    model = qc.CostModel.create(seed=3) if order == "cost" else None
    print(f"[gpu] batch {w.name} ({len(gs)} instances) {order}", flush=True)
    res = qc.mine_batch(gs, [g.gamma for g in gs], [g.min_size for g in gs], windows=[w.window] * len(gs),
                        order=order, cost_model=model, replays=2)
    for r, gold in zip(res, golds):
        _check(r, gold)


@pytest.mark.parametrize("wl", BATCHES)
def test_batch_workload_root_cut(wl, monkeypatch):
    """The batch in one launch with every instance's leading root removals
    applied before it (forced: these roots are below the default threshold);
    each instance's counters must still equal its own oracle result."""
    w, gs, golds = _batch(wl)
    monkeypatch.setenv("QCMINE_ROOT_CUT_WORDS", "0")
    print(f"[gpu] batch {w.name} ({len(gs)} instances), root cut forced", flush=True)
    res = qc.mine_batch(gs, [g.gamma for g in gs], [g.min_size for g in gs], windows=[w.window] * len(gs))
    for r, gold in zip(res, golds):
        _check(r, gold)


@pytest.mark.parametrize("wl", BATCHES)
def test_batch_workload_plain_subset(wl):
    """Every 16th instance on its own in the plain level-synchronous form."""
    w, gs, golds = _batch(wl)
    for i in range(0, len(gs), 16):
        print(f"[gpu] batch {w.name} instance {i} plain", flush=True)
        r = qc.mine(gs[i], gs[i].gamma, gs[i].min_size, 0, mode="plain", window=w.window)
        _check(r, golds[i])
