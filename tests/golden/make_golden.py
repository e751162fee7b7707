"""Regenerate the golden fixtures in this folder with the CPU oracle.

    python tests/golden/make_golden.py [--windows 1,2] [--leak-windows 1,2]

workload0.json      generator checksum, oracle counters and maximal sets of
                    workload 0 (full search)
small_vectors.json  seeded small graphs (CSR inline) with task counts and
                    maximal sets
map_leak.json       faithful against clean map mode (the reference's stale map
                    entries, main.cu:1694-1697): cases whose raw output
                    differs, and stale positions per case (--leak-windows
                    adds workload windows)
windows.json        (--windows) per-workload window results: counters and the
                    SHA-256 of the maximal-set file, computed by running the
                    oracle over disjoint root-branch ranges in parallel
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests.graphs import SMALL_CASES, planted  # noqa: E402

STATS = ["tasks", "branches", "children", "raw_cliques", "w_adj", "sum_t", "sum_t_written", "sum_clique_len",
         "beyond"]


def digest(sets) -> str:
    h = hashlib.sha256()
    for s in sets:
        h.update((" ".join(str(int(x)) for x in s) + "\n").encode())
    return h.hexdigest()


def workload0():
    from cuqc_amd import Graph
    g = Graph.generate(0, 1)
    off, adj = g.csr()
    r = oracle.mine(off, adj, g.gamma, g.min_size)
    mx = oracle.nonmax(r["raw"])
    return {"checksum": hex(g.checksum()), "gamma": g.gamma, "min_size": g.min_size,
            "stats": {k: r["stats"][k] for k in STATS}, "maximal": [list(s) for s in mx],
            "maximal_sha256": digest(mx)}


def small_vectors():
    out = []
    specs = [(f"planted{i}", 24 + 4 * i, 0.12, [8 + i, 10], 0.9, 0.8, 6) for i in range(6)]
    specs += [(f"dense{i}", 18 + 2 * i, 0.5, [], 1.0, 0.7, 5) for i in range(6)]
    for i, (name, n, p, sizes, keep, gamma, ms) in enumerate(specs):
        off, adj = planted(n, p, sizes, keep, 1000 + i)
        r = oracle.mine(off, adj, gamma, ms)
        out.append({"name": name, "off": [int(x) for x in off], "adj": [int(x) for x in adj], "gamma": gamma,
                    "min_size": ms, "tasks": r["stats"]["tasks"],
                    "maximal": [list(s) for s in oracle.nonmax(r["raw"])]})
    return out


def _leak_case(args):
    """Faithful (stale map entries carried over, main.cu:1694-1697) against
    clean map mode on one graph and window: do the raw sets or the task
    count differ, and did any stale position fall outside its record."""
    name, off, adj, gamma, ms, win = args
    kw = dict(root_begin=win[0], root_end=win[1], max_depth=win[2]) if win else {}
    a = oracle.mine(off, adj, gamma, ms, faithful=False, **kw)
    b = oracle.mine(off, adj, gamma, ms, faithful=True, **kw)
    differs = oracle.canon(a["raw"]) != oracle.canon(b["raw"]) or a["stats"]["tasks"] != b["stats"]["tasks"]
    return name, differs, b["stats"]["stale_oob"], b["stats"]["stale_hits"]


def map_leak(windows=()):
    from cuqc_amd import Graph, workloads
    g = Graph.generate(0, 1)
    cases = [("config0", *g.csr(), g.gamma, g.min_size, None)]
    cases += [(n, *b(), gm, ms, None) for n, b, gm, ms in SMALL_CASES]
    for wl in windows:
        w = workloads.WORKLOADS[wl]
        gw = workloads.graph(w)
        cases.append((w.label, *gw.csr(), gw.gamma, gw.min_size, w.window))
    out = {"differs": [], "stale_oob": {}, "stale_hits": {}}
    with ProcessPoolExecutor(int(os.environ.get("GOLDEN_JOBS", 8))) as ex:
        for name, differs, oob, hits in ex.map(_leak_case, cases):
            if differs:
                out["differs"].append(name)
            out["stale_oob"][name] = oob
            out["stale_hits"][name] = hits
    return out


def _window_part(args):
    """Counters of root branches [b0, b1) alone: the run over [b0, b1) minus
    the run over the empty window [b0, b0), which holds the root's own work
    (its task, lookahead and the removals that lead up to branch b0); the
    first part of a window also returns that root share."""
    wl, inst, b0, b1, first, threads = args
    from cuqc_amd import workloads
    w = workloads.WORKLOADS[wl]
    g = workloads.graph(w, inst)
    off, adj = g.csr()
    r = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=b0, root_end=b1, max_depth=w.max_depth,
                    threads=threads)
    e = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=b0, root_end=b0, max_depth=w.max_depth)
    st = {k: r["stats"][k] - e["stats"][k] for k in STATS}
    return st, [[int(x) for x in s] for s in r["raw"]], e["stats"] if first else None


def window_result(wl: int, inst: int = 0, jobs: int = 8, parts: int = 64) -> dict:
    from cuqc_amd import workloads
    w = workloads.WORKLOADS[wl]
    g = workloads.graph(w, inst)
    off, adj = g.csr()
    nb = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=0, root_end=0)["stats"]["root_branches"]
    # a negative window start counts back from the last root branch
    b0 = w.root_begin if w.root_begin >= 0 else max(0, nb + w.root_begin)
    b1 = nb if w.root_end < 0 else min(nb, w.root_end)
    cuts = sorted(set([b0] + [b0 + (b1 - b0) * i // parts for i in range(1, parts)] + [b1]))
    # fewer parts than workers (a window of a few deep root branches): each
    # part runs the oracle on several threads
    jobs = int(os.environ.get("GOLDEN_JOBS", jobs))
    threads = max(1, jobs // max(1, len(cuts) - 1))
    # the first part also returns the root's own counters
    tasks = [(wl, inst, a, b, a == b0, threads) for a, b in zip(cuts, cuts[1:])]
    tot = {k: 0 for k in STATS}
    raw = []
    with ProcessPoolExecutor(min(jobs, len(tasks))) as ex:
        for st, r, root in ex.map(_window_part, tasks):
            for k in STATS:
                tot[k] += st[k] + (root[k] if root is not None else 0)
            raw += r
    mx = oracle.nonmax(raw)
    return {"workload": w.name, "instance": inst, "checksum": hex(g.checksum()), "window": list(w.window),
            "root_branches": nb, "stats": tot, "raw_sha256": digest(oracle.canon(raw)), "maximal": len(mx),
            "maximal_sha256": digest(mx)}


def _instance_result(args):
    wl, inst = args
    from cuqc_amd import workloads
    w = workloads.WORKLOADS[wl]
    g = workloads.graph(w, inst)
    off, adj = g.csr()
    r = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=w.root_begin, root_end=w.root_end,
                    max_depth=w.max_depth)
    mx = oracle.nonmax(r["raw"])
    return {"workload": w.name, "instance": inst, "checksum": hex(g.checksum()), "window": list(w.window),
            "gamma": g.gamma, "min_size": g.min_size, "root_branches": r["stats"]["root_branches"],
            "stats": {k: r["stats"][k] for k in STATS}, "raw_sha256": digest(oracle.canon(r["raw"])),
            "maximal": len(mx), "maximal_sha256": digest(mx)}


def batch_results(wl: int, jobs: int = 8) -> dict:
    """Per-instance oracle results of a batch workload (one process per
    instance)."""
    from cuqc_amd import workloads
    w = workloads.WORKLOADS[wl]
    out = {}
    with ProcessPoolExecutor(int(os.environ.get("GOLDEN_JOBS", jobs))) as ex:
        for r in ex.map(_instance_result, [(wl, i) for i in range(w.instances)]):
            out[f"{wl}:{r['instance']}"] = r
            print(f"[golden] {wl}:{r['instance']} tasks {r['stats']['tasks']}", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--windows", default="")
    ap.add_argument("--leak-windows", default="",
                    help="workloads whose windows map_leak.json also covers (single-threaded: slow)")
    a = ap.parse_args()
    with open(os.path.join(HERE, "workload0.json"), "w") as f:
        json.dump(workload0(), f)
    with open(os.path.join(HERE, "small_vectors.json"), "w") as f:
        json.dump(small_vectors(), f)
    if a.leak_windows:
        res = map_leak([int(x) for x in a.leak_windows.split(",") if x])
        with open(os.path.join(HERE, "map_leak.json"), "w") as f:
            json.dump(res, f, indent=1)
    if a.windows:
        path = os.path.join(HERE, "windows.json")
        cur = json.load(open(path)) if os.path.exists(path) else {}
        from cuqc_amd import workloads
        for wl in [int(x) for x in a.windows.split(",")]:
            if workloads.WORKLOADS[wl].instances > 1:
                cur.update(batch_results(wl))
            else:
                cur[str(wl)] = window_result(wl, parts=int(os.environ.get("GOLDEN_PARTS", 64)))
                print(json.dumps(cur[str(wl)]), flush=True)
            # saved after every workload
            with open(path, "w") as f:
                json.dump(cur, f, indent=1)


if __name__ == "__main__":
    main()
