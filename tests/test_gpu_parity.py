"""GPU path vs the CPU oracle: maximal sets bit-exact, raw-set multiset and
task counters identical (every task record is processed as a pure function
of its record, so both are scheduling-invariant)."""
import sys

import numpy as np
import pytest

import cuqc_amd as qc
import oracle
from tests.graphs import SMALL_CASES

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(res, ref, label):
    assert res.stats["tasks"] == ref["stats"]["tasks"], label
    assert res.stats["branches"] == ref["stats"]["branches"], label
    assert res.stats["children"] == ref["stats"]["children"], label
    assert res.stats["raw_cliques"] == ref["stats"]["raw_cliques"], label
    assert res.stats["w_adj"] == ref["stats"]["w_adj"], label
    assert res.stats["sum_t"] == ref["stats"]["sum_t"], label
    assert res.stats["sum_t_written"] == ref["stats"]["sum_t_written"], label
    assert oracle.canon(res.raw) == oracle.canon(ref["raw"]), label
    assert res.maximal_sorted_tuples() == oracle.nonmax(ref["raw"]), label


@pytest.fixture(scope="module")
def cfg0():
    g = qc.Graph.generate(0, 1)
    off, adj = g.csr()
    return g, oracle.mine(off, adj, g.gamma, g.min_size)


@pytest.mark.parametrize("lds_tile", [64, 128, 256])
@pytest.mark.parametrize("sched", [0, 1])
def test_config0_plain(cfg0, lds_tile, sched):
    g, ref = cfg0
    _say(f"[gpu] config0 plain lds_tile={lds_tile} sched={sched}")
    res = qc.mine(g, g.gamma, g.min_size, sched, mode="plain", lds_tile=lds_tile)
    _compare(res, ref, f"config0 tile {lds_tile} sched {sched}")
    assert len(res.maximal) == 3


@pytest.mark.parametrize("name,build,gamma,ms", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_small_plain(name, build, gamma, ms):
    off, adj = build()
    _say(f"[gpu] small {name} n={len(off) - 1}")
    g = qc.Graph.from_csr(off, adj)
    ref = oracle.mine(off, adj, gamma, ms)
    for tile in (64, 128):
        res = qc.mine(g, gamma, ms, 0, mode="plain", lds_tile=tile)
        _compare(res, ref, f"{name} tile {tile}")


@pytest.mark.parametrize("lds_tile", [64, 128])
def test_config0_persistent(cfg0, lds_tile):
    g, ref = cfg0
    _say(f"[gpu] config0 persistent lds_tile={lds_tile}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=lds_tile)
    _compare(res, ref, f"config0 persistent tile {lds_tile}")


@pytest.mark.parametrize("name,build,gamma,ms", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_small_persistent(name, build, gamma, ms):
    off, adj = build()
    _say(f"[gpu] small persistent {name} n={len(off) - 1}")
    g = qc.Graph.from_csr(off, adj)
    ref = oracle.mine(off, adj, gamma, ms)
    res = qc.mine(g, gamma, ms, 0, mode="persistent", lds_tile=64)
    _compare(res, ref, f"{name} persistent")


@pytest.mark.parametrize("window", [(0, 40, -1), (40, 100, -1), (0, -1, 2), (10, 60, 3)])
def test_config0_window(cfg0, window):
    g, _ = cfg0
    off, adj = g.csr()
    _say(f"[gpu] config0 window {window}")
    ref = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=window[0], root_end=window[1], max_depth=window[2])
    for mode in ("plain", "persistent"):
        res = qc.mine(g, g.gamma, g.min_size, 0, mode=mode, window=window)
        _compare(res, ref, f"config0 window {window} {mode}")
        assert res.stats["beyond"] == ref["stats"]["beyond"]


@pytest.mark.parametrize("window", [(40, 100, -1), (10, 60, 3), (150, -1, -1), (-20, -1, 2), (5, 6, -1)])
def test_config0_window_root_cut(cfg0, window, monkeypatch):
    """A window's leading removals applied to the root before the launch
    (forced here for a small root) give the oracle's counters and sets."""
    g, _ = cfg0
    off, adj = g.csr()
    monkeypatch.setenv("QCMINE_ROOT_CUT_WORDS", "0")
    _say(f"[gpu] config0 window {window}, root cut forced")
    ref = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=window[0], root_end=window[1], max_depth=window[2])
    for mode in ("plain", "persistent"):
        res = qc.mine(g, g.gamma, g.min_size, 0, mode=mode, window=window)
        _compare(res, ref, f"config0 window {window} {mode} cut")
        assert res.stats["beyond"] == ref["stats"]["beyond"]


def test_root_cut_skipped_when_the_root_is_a_quasi_clique(monkeypatch):
    """K_30: the whole root passes the lookahead and is the one maximal set;
    a window starting past branch 0 must not shorten it."""
    n = 30
    off = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.uint64)
    adj = np.array([u for v in range(n) for u in range(n) if u != v], np.uint32)
    g = qc.Graph.from_csr(off, adj)
    monkeypatch.setenv("QCMINE_ROOT_CUT_WORDS", "0")
    ref = oracle.mine(off, adj, 0.9, 10, root_begin=5, root_end=20)
    _say("[gpu] K_30 window (5, 20), root cut forced")
    for mode in ("plain", "persistent"):
        res = qc.mine(g, 0.9, 10, 0, mode=mode, window=(5, 20, -1))
        _compare(res, ref, f"K30 window {mode}")
    assert oracle.nonmax(ref["raw"]) == [tuple(range(n))]


@pytest.mark.parametrize("chunk", [1, 3, 16])
def test_config0_persistent_split(cfg0, chunk):
    g, ref = cfg0
    _say(f"[gpu] config0 persistent split chunk={chunk}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=64, split_chunk=chunk, split_min=1)
    _compare(res, ref, f"config0 split {chunk}")


@pytest.mark.parametrize("order", ["fifo", "cost"])
@pytest.mark.parametrize("workgroups", [4, 8])
def test_config0_deep_queue_multi_claims(cfg0, order, workgroups):
    """Few resident waves and one-branch items keep the queue many times
    deeper than the grid, so waves claim 2 or 4 consecutive slots at once."""
    g, ref = cfg0
    _say(f"[gpu] config0 deep queue order={order} workgroups={workgroups}")
    kw = dict(order="cost", cost_model=qc.CostModel.create(seed=3)) if order == "cost" else {}
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=64, split_chunk=1, split_min=1,
                  workgroups=workgroups, **kw)
    _compare(res, ref, f"config0 deep queue {order} {workgroups}")


@pytest.mark.parametrize("name,build,gamma,ms", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_small_persistent_split(name, build, gamma, ms):
    off, adj = build()
    _say(f"[gpu] small persistent split {name}")
    g = qc.Graph.from_csr(off, adj)
    ref = oracle.mine(off, adj, gamma, ms)
    res = qc.mine(g, gamma, ms, 0, mode="persistent", lds_tile=64, split_chunk=2, split_min=1)
    _compare(res, ref, f"{name} split")


@pytest.mark.parametrize("window", [(5, 90, -1), (0, -1, 2)])
def test_config0_window_split(cfg0, window):
    g, _ = cfg0
    off, adj = g.csr()
    _say(f"[gpu] config0 window split {window}")
    ref = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=window[0], root_end=window[1], max_depth=window[2])
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=64, split_chunk=4, split_min=1, window=window)
    _compare(res, ref, f"config0 window {window} split")


def test_batched_small_instances():
    """Many instances through one persistent launch: each instance's result
    equals its own oracle run."""
    cases = [(n, *b(), gm, ms) for n, b, gm, ms in SMALL_CASES]
    g0 = qc.Graph.generate(0, 1)
    cases.append(("config0", *g0.csr(), g0.gamma, g0.min_size))
    _say(f"[gpu] batched {len(cases)} instances")
    graphs = [qc.Graph.from_csr(off, adj) for _, off, adj, _, _ in cases]
    res = qc.mine_batch(graphs, [c[3] for c in cases], [c[4] for c in cases], lds_tile=64, split_chunk=4, split_min=1)
    for (name, off, adj, gamma, ms), r in zip(cases, res):
        _compare(r, oracle.mine(off, adj, gamma, ms), f"batched {name}")


def test_batched_graph_replays():
    g0 = qc.Graph.generate(0, 1)
    off, adj = g0.csr()
    ref = oracle.mine(off, adj, g0.gamma, g0.min_size)
    _say("[gpu] batched graph replays x4")
    res = qc.mine_batch([g0, g0], [g0.gamma] * 2, [g0.min_size] * 2, lds_tile=128, replays=4)
    for r in res:
        _compare(r, ref, "replays")


def _trace_rows(feats, subtree):
    return sorted((tuple(float(x) for x in f), int(s)) for f, s in zip(feats, subtree))


@pytest.mark.parametrize("chunk,window", [(-1, None), (4, None), (-1, (0, -1, 2)), (3, (5, -1, 3))])
def test_trace_matches_oracle(cfg0, chunk, window):
    """The GPU task trace (features + subtree sizes, children past a depth
    window included) equals the oracle's."""
    from cuqc_amd import costmodel as cm
    g, _ = cfg0
    off, adj = g.csr()
    w = window or (0, -1, -1)
    ref = oracle.mine(off, adj, g.gamma, g.min_size, trace=True, root_begin=w[0], root_end=w[1], max_depth=w[2])
    sub = cm.subtree_sizes(ref["trace_parent"], ref["trace_beyond"])
    _say(f"[gpu] trace chunk={chunk} window={window}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=64, split_chunk=chunk, split_min=1, trace=True,
                  window=window)
    assert len(res.trace_subtree) == ref["stats"]["tasks"]
    assert _trace_rows(res.trace_feats, res.trace_subtree) == _trace_rows(ref["trace_feats"], sub)


def test_capacity_growth_reruns(cfg0, monkeypatch):
    """Capacities start small and grow by rerunning the launch: same result."""
    g, ref = cfg0
    monkeypatch.setenv("QCMINE_INITIAL_CAPACITY", "64")
    _say("[gpu] capacity growth from 64")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", split_chunk=4, split_min=1)
    _compare(res, ref, "growth")


@pytest.mark.parametrize("mode", ["plain", "persistent"])
def test_fixed_capacity_overflow_is_an_error(cfg0, mode):
    g, _ = cfg0
    _say(f"[gpu] fixed capacity overflow {mode}")
    with pytest.raises(qc.QCError) as e:
        qc.mine(g, g.gamma, g.min_size, 0, mode=mode, task_capacity=300, clique_capacity=50)
    assert e.value.code == qc.QC_EOVERFLOW


@pytest.mark.parametrize("lds_wide", [-1, 128, 1024])
def test_config0_lds_wide(cfg0, lds_wide):
    """Tasks above the 64-vertex LDS tile with their arrays in LDS (or not)."""
    g, ref = cfg0
    _say(f"[gpu] config0 lds_wide={lds_wide}")
    res = qc.mine(g, g.gamma, g.min_size, 0, mode="persistent", lds_tile=64, lds_wide=lds_wide, split_chunk=3, split_min=1)
    _compare(res, ref, f"lds_wide {lds_wide}")


@pytest.mark.parametrize("name,build,gamma,ms", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_small_lds_wide(name, build, gamma, ms):
    off, adj = build()
    ref = oracle.mine(off, adj, gamma, ms)
    g = qc.Graph.from_csr(off, adj)
    _say(f"[gpu] small lds_wide {name}")
    res = qc.mine(g, gamma, ms, 0, mode="persistent", lds_tile=64, lds_wide=256, split_chunk=2, split_min=1)
    _compare(res, ref, name)


def test_concurrent_calls_on_two_threads(cfg0):
    """Two host threads mine at once (the library's per-device block cache and
    workspace are shared between them): each result equals the oracle's.  One
    thread runs batches that hold the same graph twice (its device copy is
    made once, under a lock)."""
    import threading
    g, ref = cfg0
    cases = [(n, *b(), gm, ms) for n, b, gm, ms in SMALL_CASES[:4]]
    refs = [oracle.mine(off, adj, gm, ms) for _, off, adj, gm, ms in cases]
    errors = []

    def plain_worker():
        try:
            for it in range(3):
                for (name, off, adj, gm, ms), r in zip(cases, refs):
                    res = qc.mine(qc.Graph.from_csr(off, adj), gm, ms, 0, mode="plain")
                    _compare(res, r, f"thread plain {name} #{it}")
        except Exception as e:  # reported by the main thread
            errors.append(e)

    def batch_worker():
        try:
            for it in range(3):
                gg = qc.Graph.generate(0, 1)
                out = qc.mine_batch([gg, gg], [gg.gamma] * 2, [gg.min_size] * 2, split_chunk=4, split_min=1)
                for r in out:
                    _compare(r, ref, f"thread batch #{it}")
        except Exception as e:
            errors.append(e)

    _say("[gpu] two threads: plain calls and batches at once")
    ts = [threading.Thread(target=plain_worker), threading.Thread(target=batch_worker)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts), "a mining thread did not finish"
    if errors:
        raise errors[0]
