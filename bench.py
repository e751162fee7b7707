#!/usr/bin/env python3
"""Enumeration throughput of the qcmine GPU path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1]

A step is one enumeration of the workload (root construction, expansion to
completion inside the workload's search window, raw-clique collection) with
the graph already resident in HBM; RemoveNonMax runs after the timed window
(the reference times it separately, main.cu:497-509).  The JSON line reports
tasks/s of the fastest GPU form, the per-form enumeration times, the roofline
of the expansion kernels and the CPU oracle's throughput on a bounded sample
of the same workload.

Algorithmic bytes per step (the roofline numerator), counting only work
that is done inside the window:
    A = 12 * sum_t + 12 * (sum_t - root_size) + 4 * sum_clique_len
(12 B per task vertex read, and written once by the parent: every expanded
task but the root was written; children past a depth window are built but
never written and are not counted; 4 B per emitted clique vertex).  The
reference's adjacency-list scans (4 B per entry, w_adj entries) are reported
beside it as list_scan_bytes but not counted: here every such scan is a test
of a bit row of the root's 1-hop / 2-hop matrices, which stay in L2 and LDS,
and on the dense config-1 window the list volume (about 300 GB per step)
exceeds what HBM could deliver in the kernel's time by two orders of
magnitude.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel (expand_persistent) in the
# headline form, keyed by workload label: one rocprofv3 FETCH_SIZE and one
# WRITE_SIZE pass, each scaled by the counter's reading for a 256 MiB device
# copy (FETCH x2.000, WRITE x1.000; profiles/expand_persistent_*.md).  The
# autotuner may pick another schedule on another run; the figure belongs to
# the profiled one.  None where not measured.
TRAFFIC: dict = {
    # 739,946 KiB x 2.000 + 21,237,700 KiB x 1.000
    "coexpr8k[3600:3612,depth<=3]": 739946 * 1024 * 2 + 21237700 * 1024,
}


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def set_digest(sets) -> str:
    h = hashlib.sha256()
    for s in sets:
        h.update((" ".join(str(int(x)) for x in s) + "\n").encode())
    return h.hexdigest()


def algo_bytes(st: dict) -> int:
    return 12 * st["sum_t"] + 12 * (st["sum_t"] - st["root_size"]) + 4 * st["sum_clique_len"]


def cpu_info() -> dict:
    """Host CPU model (lscpu / /proc/cpuinfo) and the cores this process may use."""
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"cpu_model": model, "nproc": nproc}


def cpu_baseline(w, budget_s: float) -> dict:
    """Time the oracle on a bounded sample of the workload: its last k root
    branches, with the depth limit cut to at most 2 (a single root branch of
    the config-1 window at depth 3 is half a minute of one core), k doubled
    until the run takes about budget_s / 2 seconds or more.  One thread (the
    reference's host path has none), and the same sample on every host core
    as "cpu_best" (not the reference: the oracle's threaded search)."""
    import oracle
    from cuqc_amd import workloads

    g = workloads.graph(w)
    off, adj = g.csr()
    b1 = w.root_end
    if b1 < 0:
        b1 = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=0, root_end=0)["stats"]["root_branches"]
    lo = w.root_begin if w.root_begin >= 0 else 0
    depth = 2 if w.max_depth < 0 else min(w.max_depth, 2)
    k, best = 1, None
    t_spent = 0.0
    while True:
        b0 = max(lo, b1 - k)
        t0 = time.perf_counter()
        r = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=b0, root_end=b1, max_depth=depth)
        dt = time.perf_counter() - t0
        t_spent += dt
        best = (r["stats"]["tasks"], dt, b0, b1)
        if dt >= budget_s / 2 or b0 == lo or t_spent > budget_s:
            break
        k *= 4 if dt < budget_s / 64 else 2
    tasks, dt, b0, b1 = best
    info = cpu_info()
    nthr = max(1, min(int(info["nproc"] or 1), 64))
    t0 = time.perf_counter()
    rm = oracle.mine(off, adj, g.gamma, g.min_size, root_begin=b0, root_end=b1, max_depth=depth, threads=nthr)
    dtm = time.perf_counter() - t0
    sample = (f"oracle (C++ restatement of main.cu host path), root branches [{b0},{b1}) of {w.name} at "
              f"depth <= {depth}: {tasks} tasks")
    return {"value": round(tasks / dt, 1), "unit": "tasks/s", "cores": 1, "kind": "port",
            "sample": f"{sample} in {dt:.2f} s on one thread", **info,
            "cpu_best": {"value": round(rm["stats"]["tasks"] / dtm, 1), "unit": "tasks/s", "cores": nthr,
                         "kind": "port", "sample": f"{sample} in {dtm:.2f} s on {nthr} threads (threaded oracle)"}}


def cost_model_jobs():
    """Training traces for the cost model, from configs 1-3 and disjoint from
    every measured window: workload 0 in full, config 1 root branches
    3580-3599 (depth <= 2), config 2 root branches 9255-9262 (full depth),
    and 8 config-3 instances drawn with another seed (44) in the batch's
    window shape (last 32 root branches, depth <= 2)."""
    from cuqc_amd import workloads
    g0 = workloads.graph(workloads.WORKLOADS[0])
    jobs = [(g0, g0.gamma, g0.min_size, None)]
    g1 = workloads.graph(workloads.WORKLOADS[1])
    jobs.append((g1, g1.gamma, g1.min_size, (3580, 3600, 2)))
    g2 = workloads.graph(workloads.WORKLOADS[2])
    jobs.append((g2, g2.gamma, g2.min_size, (9255, 9263, -1)))
    from cuqc_amd import Graph
    for i in range(8):
        g3 = Graph.generate(3, 44, i)
        jobs.append((g3, g3.gamma, g3.min_size, workloads.WORKLOADS[3].window))
    return jobs


class Job:
    """The graphs of one workload and how each form mines them: one call per
    graph for the plain form, one batched persistent launch otherwise."""

    def __init__(self, qc, w):
        from cuqc_amd import workloads
        self.qc, self.w = qc, w
        self.gs = workloads.graphs(w)
        self.gammas = [g.gamma for g in self.gs]
        self.min_sizes = [g.min_size for g in self.gs]
        self.windows = [w.window] * len(self.gs)

    def run(self, kw: dict) -> dict:
        qc = self.qc
        if kw.get("mode") == "plain" or len(self.gs) == 1:
            res = [qc.mine(g, g.gamma, g.min_size, 0, window=self.w.window, **kw) for g in self.gs]
            enum_ms = sum(r.stats["t_enum_ms"] for r in res)
            kern_ms = sum(r.stats["t_kernel_ms"] for r in res)
        else:
            kb = {k: v for k, v in kw.items() if k != "mode"}
            res = qc.mine_batch(self.gs, self.gammas, self.min_sizes, windows=self.windows, **kb)
            enum_ms = res[0].stats["t_enum_ms"]
            kern_ms = res[0].stats["t_kernel_ms"]
        st = {k: sum(r.stats[k] for r in res) for k in ("tasks", "raw_cliques", "sum_t", "sum_t_written",
                                                        "sum_clique_len", "w_adj", "root_size", "beyond")}
        h = hashlib.sha256()
        for r in res:
            h.update(set_digest(r.maximal).encode())
        return dict(res=res, stats=st, enum_ms=enum_ms, kern_ms=kern_ms, digest=h.hexdigest(),
                    maximal=sum(len(r.maximal) for r in res))


SCHEDULE_CACHE = None
_MODEL = {}


def trained_model():
    """The cost model, trained once per run on cost_model_jobs() (untimed)."""
    if "m" not in _MODEL:
        from cuqc_amd import costmodel
        t0 = time.perf_counter()
        model, rows = costmodel.train_from_traces(cost_model_jobs())
        _MODEL["m"] = model
        _MODEL["info"] = {"trace_rows": rows, "seconds": round(time.perf_counter() - t0, 2),
                          "traces": "workload 0; coexpr8k[3580:3600,depth<=2]; chunglu100k[9255:9263,full depth]; "
                                    "8 config-3 graphs of seed 44 [-32:,depth<=2]"}
    return _MODEL["m"], _MODEL["info"]


def setup_forms(job: Job, forms, tune: bool = True):
    """Untimed set-up: the trained cost model and the autotuned schedules, each
    tuned in the order its form runs (FIFO for `tuned`; cost order for `cost`
    when `tune`: the best schedule of one order need not be the other's)."""
    from cuqc_amd import autotune
    info, kws = {}, {"plain": dict(mode="plain"), "persistent": dict(mode="persistent")}

    def tuned(key: str, **order_kw) -> dict:
        t0 = time.perf_counter()
        sched = autotune.tune(job.gs, job.gammas, job.min_sizes, job.windows, reps=1, log=_log, cache=SCHEDULE_CACHE,
                              **order_kw)
        info[key] = dict(sched.kwargs(), trials=len(sched.trials), seconds=round(time.perf_counter() - t0, 2))
        return sched.kwargs()

    if "tuned" in forms:
        kws["tuned"] = dict(mode="persistent", **tuned("autotune_fifo"))
    if "cost" in forms:
        model, minfo = trained_model()
        info["cost_model"] = minfo
        sched_kw = tuned("autotune", order="cost", cost_model=model) if tune else {}
        kws["cost"] = dict(mode="persistent", order="cost", cost_model=model, **sched_kw)
    return kws, info


def run_form(job: Job, form: str, kw: dict, steps: int, warmup: int) -> dict:
    for i in range(warmup):
        o = job.run(kw)
        _log(f"[bench] {job.w.name} {form} warmup {i + 1}/{warmup}: enum {o['enum_ms']:.1f} ms")
    runs = []
    for i in range(steps):
        t0 = time.perf_counter()
        o = job.run(kw)
        o["wall_ms"] = (time.perf_counter() - t0) * 1e3
        runs.append(o)
        _log(f"[bench] {job.w.name} {form} step {i + 1}/{steps}: enum {o['enum_ms']:.1f} ms, "
             f"kernel {o['kern_ms']:.1f} ms, {o['stats']['tasks']} tasks")
    last = runs[-1]
    return dict(last, enum_ms=sum(r["enum_ms"] for r in runs) / steps, kern_ms=sum(r["kern_ms"] for r in runs) / steps,
                wall_ms=sum(r["wall_ms"] for r in runs) / steps)


def measure(job: Job, forms, steps: int, warmup: int, plain_steps: int, tune: bool = True):
    kws, setup = setup_forms(job, forms, tune)
    out, digest = {}, None
    for f in forms:
        # the plain form is the comparison point: no warmup
        st_, wu_ = (plain_steps, 0) if f == "plain" else (steps, warmup)
        o = run_form(job, f, kws[f], st_, wu_)
        if digest is not None and o["digest"] != digest:
            raise SystemExit(f"maximal sets differ between forms ({f})")
        digest = o["digest"]
        out[f] = o
    table = {f: {"enum_ms": round(o["enum_ms"], 3), "wall_ms": round(o["wall_ms"], 3),
                 "kernel_ms": round(o["kern_ms"], 3),
                 "tasks_per_s": round(o["stats"]["tasks"] / (o["enum_ms"] / 1e3), 1)} for f, o in out.items()}
    if "plain" in out:
        for f in out:
            if f != "plain":
                table[f]["speedup_vs_plain"] = round(out["plain"]["enum_ms"] / out[f]["enum_ms"], 3)
    return out, table, setup, digest


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--forms", default="plain,persistent,tuned,cost")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--batch-workload", type=int, default=3,
                    help="also measure this batch workload (one launch vs one plain run per instance); -1: skip")
    ap.add_argument("--batch-forms", default="plain,persistent,cost")
    ap.add_argument("--extra-workloads", default="2,4,5",
                    help="also measure these workloads (plain, persistent and cost-ordered persistent forms)")
    ap.add_argument("--extra-forms", default="plain,persistent,cost")
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--schedule-cache", default=None,
                    help="JSON file of autotuned schedules to reuse (and extend)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        x = torch.ones(a.copy_probe, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.copy_probe}), flush=True)
        return 0
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the miner runs on a single device and has no multi-GPU path "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        return 2

    import cuqc_amd as qc
    from cuqc_amd import workloads

    global SCHEDULE_CACHE
    SCHEDULE_CACHE = a.schedule_cache
    w = workloads.WORKLOADS[a.workload]
    job = Job(qc, w)
    forms = [f for f in a.forms.split(",") if f]
    out, table, setup, digest = measure(job, forms, a.steps, a.warmup, a.plain_steps)
    # the headline form: batched (single persistent launch), autotuned and
    # cost-ordered when measured, else the fastest
    best = "cost" if "cost" in out else min(forms, key=lambda f: out[f]["enum_ms"])
    st = out[best]["stats"]
    tasks = st["tasks"]
    enum_ms = out[best]["enum_ms"]
    A = algo_bytes(st)
    kern_s = out[best]["kern_ms"] / 1e3
    achieved = A / kern_s / 1e9 if kern_s > 0 else 0.0
    g = job.gs[0]
    line = {
        "metric": "enumeration ms + tasks/sec, batched+cost-ordered vs plain GPU, bit-exact set",
        "value": round(tasks / (enum_ms / 1e3), 1),
        "unit": "tasks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(enum_ms, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": w.label, "generator_config": w.config, "seed": w.seed, "instances": w.instances,
                   "gamma": g.gamma, "min_size": g.min_size, "n": g.n, "form": best},
        "forms": table,
        "tasks": tasks,
        "beyond": st["beyond"],
        "raw_cliques": st["raw_cliques"],
        "maximal_cliques": out[best]["maximal"],
        "maximal_sha256": digest,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": TRAFFIC.get(w.label),
                     "algorithmic_bytes": A, "list_scan_bytes": 4 * st["w_adj"],
                     "kernel_ms": round(out[best]["kern_ms"], 3)},
    }
    if "plain" in out and best != "plain":
        line["speedup_vs_plain"] = round(out["plain"]["enum_ms"] / enum_ms, 3)
    line["setup"] = setup
    if a.batch_workload >= 0 and a.batch_workload != a.workload:
        wb = workloads.WORKLOADS[a.batch_workload]
        bjob = Job(qc, wb)
        bforms = [f for f in a.batch_forms.split(",") if f]
        bout, btable, bsetup, bdigest = measure(bjob, bforms, a.steps, a.warmup, a.plain_steps)
        bbest = "cost" if "cost" in bout else min(bforms, key=lambda f: bout[f]["enum_ms"])
        line["batch"] = {"workload": wb.label, "instances": wb.instances, "form": bbest,
                         "tasks": bout[bbest]["stats"]["tasks"], "ms_per_step": round(bout[bbest]["enum_ms"], 3),
                         "forms": btable, "maximal_cliques": bout[bbest]["maximal"],
                         "maximal_sha256": bdigest, "setup": bsetup}
        if "plain" in bout and bbest != "plain":
            line["batch"]["speedup_vs_plain"] = round(bout["plain"]["enum_ms"] / bout[bbest]["enum_ms"], 3)
    extras = {}
    for k in [int(x) for x in a.extra_workloads.split(",") if x]:
        if k in (a.workload, a.batch_workload):
            continue
        we = workloads.WORKLOADS[k]
        ejob = Job(qc, we)
        eforms = [f for f in a.extra_forms.split(",") if f]
        eout, etable, esetup, edigest = measure(ejob, eforms, a.steps, a.warmup, a.plain_steps)
        ebest = "cost" if "cost" in eout else min(eforms, key=lambda f: eout[f]["enum_ms"])
        est = eout[ebest]["stats"]
        extras[we.label] = {"form": ebest, "tasks": est["tasks"], "beyond": est["beyond"],
                            "raw_cliques": est["raw_cliques"], "maximal_cliques": eout[ebest]["maximal"],
                            "maximal_sha256": edigest, "ms_per_step": round(eout[ebest]["enum_ms"], 3),
                            "tasks_per_s": round(est["tasks"] / (eout[ebest]["enum_ms"] / 1e3), 1),
                            "forms": etable, "setup": esetup}
        if "plain" in eout and ebest != "plain":
            extras[we.label]["speedup_vs_plain"] = round(eout["plain"]["enum_ms"] / eout[ebest]["enum_ms"], 3)
    if extras:
        line["workloads"] = extras
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, a.cpu_budget_s)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
